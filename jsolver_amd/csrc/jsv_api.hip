// Host side of the C ABI declared in include/jsolver_amd.h: plans (per-level
// device buffers), coefficient setup, the recursive V/W/F cycle and the solve
// loop.  The reference builds the same recursion at trace time in Python
// (multigrid_cycle, jsolver/solver_2d.py:582-614); here it is a C++ recursion
// that enqueues kernels on the caller's stream.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/jsolver_amd.h"
#include "jsv_kernels.hpp"

#pragma clang fp contract(off)

using jsv::Geo;

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return 1;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                         \
    } while (0)

#define LAUNCHCHK()                                                                           \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return fail("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                            \
    } while (0)

#define RETCHK(expr)          \
    do {                      \
        int r_ = (expr);      \
        if (r_ != 0) return r_; \
    } while (0)

template <typename T>
struct Hier {
    std::vector<Geo> g;
    std::vector<T*> phi, rhs, cx, cy, cz, cc, drr, dlr;
    std::vector<T*> res;   // first-half residual scratch of the two-pass restriction (large levels)
    std::vector<T*> phi2;  // second phi buffer of the fused out-of-place 3D sweep (large levels)
};

inline dim3 grid_points(const Geo& g, int dim, int xcount)
{
    // xcount points along x starting at 2, my-1 rows, mz-1 planes (3D) or
    // one z per batched problem (2D)
    return dim3((unsigned)((xcount + jsv::BX - 1) / jsv::BX),
                (unsigned)((g.my - 1 + jsv::BY - 1) / jsv::BY),
                (unsigned)(dim == 3 ? g.mz - 1 : g.nz));
}
// one thread per element, for kernels without a grid-stride loop
inline dim3 grid_exact(long n)
{
    long b = (n + 255) / 256;
    return dim3((unsigned)(b < 1 ? 1 : b));
}
// capped grid for grid-stride kernels
inline dim3 grid_flat(long n)
{
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

}  // namespace

constexpr int ACTK_ROWS = 64;        // cycles per staged block of per-problem activity rows

struct jsv_plan {
    int dim = 0, prec = 0, levels = 0;
    int batch = 1;                  // 2D: independent problems solved together
    int m[3] = {0, 0, 0};          // x, y, z resolutions of level 0
    double idel[3] = {0, 0, 0};
    bool op_set = false, spatial = false, offdiag = false;
    bool lam_array = false;         // lambda set from an array (else the scalar)
    Hier<double> h64;
    Hier<float> h32;
    double* phi_old = nullptr;      // level-0 copy for the distance
    double* partials = nullptr;
    long npartials = 0;
    double* dscal = nullptr;        // device scalars: 4 slots x batch
    double* hscal = nullptr;        // pinned host mirror
    int* d_act = nullptr;           // per-problem "still iterating" flags
    int* h_act = nullptr;
    int* d_actk = nullptr;          // per-cycle activity rows of a cycle block (ACTK_ROWS x batch)
    int* h_actk = nullptr;          // pinned staging of d_actk
    long slab_bytes = 0;            // slab size of the 3D two-launch smoothing schedule (0: whole level)
    long two_pass_min[2] = {1L << 20, 1L << 20};   // FP64 / FP32: level size from which restriction runs in two passes
    int fused_min_m = 256;          // 3D levels with mz >= this use the fused out-of-place sweep
    int zchunk = 0;                 // planes per workgroup of the fused sweep (0: by level size)
    int sweep_rows = 0;             // rows per workgroup tile of the fused sweep (0: by precision)
    int sweep_threads = 0;          // threads per workgroup of the fused sweep (0: 1024; FP32 two-pair
                                    // sweeps without the prolongation 768)
    int sweep_variant = 2;          // 1: thread per ring point; 2: thread per column pair, pipelined
    int sweep_pairs = 0;            // pair kernel: column pairs per thread (0: 2 in FP32, 1 in FP64)
    int n_cu = 256;                 // compute units of the device
    int zigzag = 2;                 // pair sweeps march opposite to their level's previous pair sweep (2: in
                                    // z-chunks of alternating direction; 1: the whole level, plain sweeps only)
    int fuse_resid = 1;             // the last pre-smoothing pair sweep stores the restriction's residuals
    mutable unsigned zz_dn[2] = {0u, 0u};   // FP64 / FP32: bit l = level l's last pair sweep marched downward
    int fuse_prolong = 2;           // 3D: post-smoothing pair sweeps apply the prolongation (1: FP32 only, 2: both)
    int fuse_halves = 1;            // levels smoothed one launch per half: the restriction runs the coarse
                                    // level's first half, the first post-smoothing half applies the prolongation
    bool use_graphs = true;         // replay whole cycles as captured hipGraphs
    int coarse_m[2] = {0, 0};       // FP64 / FP32: levels with max resolution <= this run as one coarse program (0: off)
    void* lv_dev[2] = {nullptr, nullptr};          // LvPtr<double> / LvPtr<float> per level
    std::vector<const void*> lv_host[2];           // pointers lv_dev holds (start level + table)
    int coarse_lds = 1;             // stage the program's coarsest levels in LDS
    int coarse_solo = jsv::COARSE_WAVE;   // ops on staged levels with at most this many points run in one wave
    int coarse_reg = 2;             // 2D runs of such ops keep their levels' points in registers (2: and
                                    // runs from a level of up to 256 points, four per lane)
    int coarse_threads = 0;   // workgroup size of the coarse program; 0: 512 for 2D FP64, else 1024
    int clds_from[2] = {0, 0}, clds_to[2] = {0, 0};   // staged level range per precision
    size_t clds_bytes[2] = {0, 0};
    void* img_dev[2] = {nullptr, nullptr};   // coefficient image of the staged levels (per problem)
    size_t img_cap[2] = {0, 0};              // its capacity in bytes
    int img_n[2] = {0, 0};                   // elements per problem
    std::vector<const void*> img_key[2];     // layout and operator version the image holds
    long op_version = 0;                     // bumped by every jsv_plan_set_operator
    struct CoarseProg {
        int prec, gamma, nu1, nu2, fcyc, start, offdiag, lds_from, solo, reg, fuse;
        unsigned* ops;              // packed op words; nullptr: too long, the cycle recurses instead
        int nops;
    };
    std::vector<CoarseProg> progs;
    hipStream_t cap_stream = nullptr;   // private stream for graph capture
    struct Graph {
        int prec, gamma, fcycle, skip_pre, nu1, nu2, use_act, stop_mode;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    bool use_act = false;           // kernels consult d_act (dynamic batched solves)
    // stop rule on the device (dynamic solves of one problem that test every cycle)
    int device_stop = 1;            // tuning: use it where it applies (0: the host reads every test)
    int stop_mode = 0;              // cycles being built: 0 none, 1 relative residual, 2 distance
    jsv::StopCtl* d_stop = nullptr; // control block
    jsv::StopCtl* h_stop = nullptr; // pinned mirror
    double* d_hist = nullptr;       // stop values (StopCtl::hist)
    long hist_cap = 0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    const int* act_ptr() const { return use_act ? d_act : nullptr; }
};

namespace {

int dalloc(jsv_plan* P, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) return fail("hipMemset failed: %s", hipGetErrorString(e));
    P->allocs.push_back(*p);
    P->bytes += (int64_t)bytes;
    return 0;
}

// frees a buffer from dalloc (nullptr: nothing); the caller has synchronised
int dfree(jsv_plan* P, void* p, size_t bytes)
{
    if (p == nullptr) return 0;
    for (size_t i = 0; i < P->allocs.size(); ++i)
        if (P->allocs[i] == p) {
            P->allocs.erase(P->allocs.begin() + (long)i);
            P->bytes -= (int64_t)bytes;
            HIPCHK(hipFree(p));
            return 0;
        }
    return fail("internal: freeing a buffer the plan does not own");
}

void drop_graphs(jsv_plan* P);

template <typename T>
int alloc_hier(jsv_plan* P, Hier<T>& H, bool with_coeffs)
{
    H.g.clear();
    const int L = P->levels;
    int lv[3];
    for (int a = 0; a < P->dim; ++a) lv[a] = (int)std::lround(std::log2((double)P->m[a])) + 1;
    for (int l = 0; l < L; ++l) {
        Geo g{};
        g.mx = 1 << (lv[0] - 1 - l);
        g.my = 1 << (lv[1] - 1 - l);
        g.mz = (P->dim == 3) ? 1 << (lv[2] - 1 - l) : 0;
        g.nx = g.mx + 3;
        g.ny = g.my + 3;
        g.nz = (P->dim == 3) ? g.mz + 3 : P->batch;
        g.sy = g.nx;
        g.sz = (long)g.nx * g.ny;
        g.n = g.sz * g.nz;
        H.g.push_back(g);
    }
    auto one = [&](std::vector<T*>& v, bool need) -> int {
        v.assign(L, nullptr);
        if (!need) return 0;
        for (int l = 0; l < L; ++l) RETCHK(dalloc(P, (void**)&v[l], sizeof(T) * H.g[l].n));
        return 0;
    };
    RETCHK(one(H.phi, true));
    RETCHK(one(H.rhs, true));
    RETCHK(one(H.cx, with_coeffs));
    RETCHK(one(H.cy, with_coeffs));
    RETCHK(one(H.cz, with_coeffs && P->dim == 3));
    RETCHK(one(H.cc, with_coeffs));
    RETCHK(one(H.drr, false));
    RETCHK(one(H.dlr, false));
    // two-pass restriction on large levels (one-pass below: there the
    // stencil re-reads hit in L2 and one launch is cheaper)
    H.res.assign(L, nullptr);
    for (int l = 0; l + 1 < L; ++l)
        if (H.g[l].n >= P->two_pass_min[sizeof(T) == 4 ? 1 : 0])
            RETCHK(dalloc(P, (void**)&H.res[l], sizeof(T) * H.g[l].n));
    H.phi2.assign(L, nullptr);
    if (P->dim == 3)
        for (int l = 0; l < L; ++l)
            if (P->fused_min_m > 0 && H.g[l].mz >= P->fused_min_m)
                RETCHK(dalloc(P, (void**)&H.phi2[l], sizeof(T) * H.g[l].n));
    return 0;
}

// ---------------------------------------------------------------------------
// elementary operations
// ---------------------------------------------------------------------------
template <typename T, int TY, int NT>
int launch_rbsweep3_t(const jsv_plan* P, Hier<T>& H, int l, hipStream_t s)
{
    const Geo& g = H.g[l];
    const int ytiles = (g.my - 1 + TY - 1) / TY;
    int zchunk = P->zchunk > 0 ? P->zchunk : (sizeof(T) == 4 ? 16 : 32);
    const int zchunks = (g.mz - 1 + zchunk - 1) / zchunk;
    const int nh = (g.nx + 1) / 2;
    const size_t lds = sizeof(T) * (size_t)(3 * (TY + 4) + 3 * (TY + 2)) * nh;
    if (lds > 160 * 1024) return fail("fused sweep tile needs %zu bytes of LDS", lds);
    hipLaunchKernelGGL((jsv::k_rbsweep3<T, TY, NT>), dim3((unsigned)(ytiles * zchunks)), dim3(NT), lds,
                       s, H.phi[l], H.phi2[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], g, zchunk,
                       ytiles, P->act_ptr());
    return 0;
}

// LDS of the pair sweep with TY rows and np pairs per thread (rings + held
// black halves, plus the coarse-plane ring when the prolongation from
// coarse level gc is fused)
// ring row width of the pair kernel instantiated for level g (its NHMAX)
int pair_nhm(const Geo& g) { return (g.nx + 1) / 2 <= 130 ? 130 : 258; }

size_t pair_sweep_lds(const Geo& g, int ty, size_t w, const Geo* gc = nullptr, int np = 1, bool res = false,
                      bool fr = false, int nt = 1024)
{
    const int nhm = pair_nhm(g);
    size_t e = (size_t)(3 * (ty + 4) + 3 * (ty + 2)) * nhm + (size_t)8 * np * nt;
    if (gc != nullptr) e += (size_t)3 * (ty / 2 + 4) * (nhm + 1);
    if (res) e += (size_t)3 * (ty + 2) * nhm + (size_t)np * nt;     // ring of phi_in's red values, held black values
    if (fr) e += (size_t)3 * (ty - 2) * nhm + (size_t)4 * np * nt;  // one plane's black values and y products; holds
    return w * e;
}

// A pre-smoothing pair sweep that also stored the first-half residuals of its
// tiles' inner rows and planes (ty > 0), and the tiling the rest follows.
struct FrGeo {
    int ty = 0, zchunk = 0, ytiles = 0, dn = 0;   // dn: the sweep marched downward
};

bool pair_sweep_rows_fit(const Geo& g, int ty, size_t w, const Geo* gc = nullptr, int np = 1)
{
    const int nhe = g.mx / 2;
    return nhe % np == 0 && (ty + 2) * (nhe / np) <= 1024 && (g.nx + 1) / 2 <= 258 &&
           pair_sweep_lds(g, ty, w, gc, np) <= 160 * 1024 && (double)g.n * w < 4294967296.0 &&
           (gc == nullptr || (double)gc->n * w < 4294967296.0);
}

bool pair_sweep_fits(const Geo& g, size_t w, const Geo* gc = nullptr) { return pair_sweep_rows_fit(g, 2, w, gc); }

template <typename T, int TY, int NHM, int NP>
int pair_sweep_per_cu(const Geo& g)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, jsv::k_rbsweep3p<T, TY, 1024, NHM, false, NP>, 1024,
                                                     pair_sweep_lds(g, TY, sizeof(T), nullptr, NP)) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu;
}

template <typename T, int TY, int NP>
int pair_sweep_per_cu(const Geo& g)
{
    return (g.nx + 1) / 2 <= 130 ? pair_sweep_per_cu<T, TY, 130, NP>(g) : pair_sweep_per_cu<T, TY, 258, NP>(g);
}

template <typename T, int NP>
int pair_sweep_per_cu(const Geo& g, int ty)
{
    switch (ty) {
    case 14: return pair_sweep_per_cu<T, 14, NP>(g);
    case 6: return pair_sweep_per_cu<T, 6, NP>(g);
    case 5: return pair_sweep_per_cu<T, 5, NP>(g);
    case 4: return pair_sweep_per_cu<T, 4, NP>(g);
    default: return pair_sweep_per_cu<T, 2, NP>(g);
    }
}

// res (FP64, no prolongation): also sum the squared residual of phi_in into
// one partial per workgroup at res, their count to *nres (0: this shape has
// no such variant, the sweep ran without it)
template <typename T, int TY, int NHM, int NP>
int launch_rbsweep3p_t(const jsv_plan* P, Hier<T>& H, int l, int zchunk, const T* pc, hipStream_t s,
                       double* res, long* nres, FrGeo* fr)
{
    if (fr != nullptr) fr->ty = 0;
    const Geo& g = H.g[l];
    constexpr int NT = 1024;
    const int nh = (g.nx + 1) / 2;
    const Geo gc = pc != nullptr ? H.g[l + 1] : g;
    if (nh > NHM || !pair_sweep_rows_fit(g, TY, sizeof(T), pc ? &gc : nullptr, NP))
        return fail("pair sweep does not fit level %d", l);
    const size_t lds = pair_sweep_lds(g, TY, sizeof(T), pc ? &gc : nullptr, NP);
    const int ytiles = (g.my - 1 + TY - 1) / TY;
    const int zchunks = (g.mz - 1 + zchunk - 1) / zchunk;
    const dim3 grid((unsigned)(ytiles * zchunks));
    // a row of pairs fills whole waves: the kernel variant with the pair
    // element picks resolved per wave at compile time, unless that variant
    // needs scratch memory (register spills cost more than the picks save)
    auto no_scratch = [](const void* k) {
        hipFuncAttributes a;
        return hipFuncGetAttributes(&a, k) == hipSuccess && a.localSizeBytes == 0;
    };
    static const bool wr_pro = no_scratch((const void*)jsv::k_rbsweep3p<T, TY, NT, NHM, true, NP, 1, true>);
    static const bool wr_plain = no_scratch((const void*)jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true>);
    const bool wr = ((g.mx / 2) / NP) % 64 == 0 && (pc != nullptr ? wr_pro : wr_plain);
    auto go = [&](auto kern, const T* pcv, size_t lds_k = 0, double* rp = nullptr, T* rfo = nullptr) {
        // zig-zag: a plain sweep (with or without the first-half residuals)
        // reverses its level's previous pair sweep; with chunks of alternating
        // direction (zig-zag 2) so does the sweep with the prolongation; the
        // sweep with the residual sum marches upward
        unsigned& last = P->zz_dn[sizeof(T) == 8 ? 0 : 1];
        const unsigned bit = 1u << l;
        const bool rev = rp == nullptr && (pcv == nullptr ? P->zigzag != 0 : P->zigzag == 2);
        const int dn = (rev && (last & bit) == 0u) ? 1 : 0;
        last = dn ? (last | bit) : (last & ~bit);
        const int alt = P->zigzag == 2 && rp == nullptr ? 2 : 0;   // chunks of alternating direction
        if (fr != nullptr) fr->dn = alt ? 0 : dn;
        hipLaunchKernelGGL(kern, grid, dim3(NT), lds_k > 0 ? lds_k : lds, s, H.phi[l], H.phi2[l], H.rhs[l],
                           H.cx[l], H.cy[l], H.cz[l], H.cc[l], g, zchunk, ytiles, pcv, gc, P->act_ptr(), rp,
                           dn | alt, rfo);
    };
    if (nres != nullptr) *nres = 0;
    if constexpr (TY >= 4) {
        // the first-half residuals of the tiles' inner rows and planes as
        // well, where that variant needs no scratch memory
        const size_t lds_f = pair_sweep_lds(g, TY, sizeof(T), nullptr, NP, false, true);
        if (fr != nullptr && pc == nullptr && res == nullptr && H.res[l] != nullptr && P->fuse_resid &&
            lds_f <= 160 * 1024) {
            static const bool wr_fr =
                no_scratch((const void*)jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true, false, true>);
            static const bool rt_fr =
                no_scratch((const void*)jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, false, false, true>);
            const bool wrf = ((g.mx / 2) / NP) % 64 == 0 && wr_fr;
            if (wrf || rt_fr) {
                wrf ? go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true, false, true>, (const T*)nullptr,
                         lds_f, nullptr, H.res[l])
                    : go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, false, false, true>, (const T*)nullptr,
                         lds_f, nullptr, H.res[l]);
                fr->ty = TY;
                fr->zchunk = zchunk;
                fr->ytiles = ytiles;
                return 0;
            }
        }
    }
    if constexpr (sizeof(T) == 8 && NP == 1 && (TY == 5 || TY == 4 || TY == 2)) {
        const size_t lds_r = pair_sweep_lds(g, TY, sizeof(T), nullptr, NP, true);
        if (res != nullptr && pc == nullptr && lds_r <= 160 * 1024 && (long)grid.x <= P->npartials) {
            static const bool wr_res = no_scratch((const void*)jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true, true>);
            const bool wrr = ((g.mx / 2) / NP) % 64 == 0 && wr_res;
            wrr ? go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true, true>, (const T*)nullptr, lds_r, res)
                : go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, false, true>, (const T*)nullptr, lds_r, res);
            *nres = (long)grid.x;
            return 0;
        }
    }
    if (pc != nullptr)
        wr ? go(jsv::k_rbsweep3p<T, TY, NT, NHM, true, NP, 1, true>, pc)
           : go(jsv::k_rbsweep3p<T, TY, NT, NHM, true, NP, 1, false>, pc);
    else
        wr ? go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, true>, (const T*)nullptr)
           : go(jsv::k_rbsweep3p<T, TY, NT, NHM, false, NP, 1, false>, (const T*)nullptr);
    return 0;
}

template <typename T, int TY, int NP>
int launch_rbsweep3p_t(const jsv_plan* P, Hier<T>& H, int l, int zchunk, const T* pc, hipStream_t s, double* res,
                       long* nres, FrGeo* fr)
{
    if ((H.g[l].nx + 1) / 2 <= 130)
        return launch_rbsweep3p_t<T, TY, 130, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    return launch_rbsweep3p_t<T, TY, 258, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
}

template <typename T, int NP>
int launch_rbsweep3p_rows(const jsv_plan* P, Hier<T>& H, int l, int rows, int zchunk, const T* pc, hipStream_t s,
                          double* res, long* nres, FrGeo* fr)
{
    switch (rows) {
    case 14: return launch_rbsweep3p_t<T, 14, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    case 6: return launch_rbsweep3p_t<T, 6, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    case 5: return launch_rbsweep3p_t<T, 5, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    case 4: return launch_rbsweep3p_t<T, 4, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    default: return launch_rbsweep3p_t<T, 2, NP>(P, H, l, zchunk, pc, s, res, nres, fr);
    }
}

// A pair sweep with NT threads per workgroup: plain, with the prolongation
// from pc, or (fr set) storing the first-half residuals (FR).  FP32 two-pair
// tiles of 10 rows take 768 threads, whose waves may hold 168 VGPRs (at 1024
// threads and 14 rows the FR variant spills).  Returns 1 (nothing launched)
// where the shape does not fit; launch_rbsweep3p picks the variants that
// run so.
template <typename T, int TY, int NHM, int NP, int NT, bool PRO, bool FR>
int launch_rbsweep3p_nt_v(const jsv_plan* P, Hier<T>& H, int l, const T* pc, hipStream_t s, FrGeo* fr)
{
    const Geo& g = H.g[l];
    const Geo gc = PRO ? H.g[l + 1] : g;
    const int nhe = g.mx / 2;
    const size_t lds = pair_sweep_lds(g, TY, sizeof(T), PRO ? &gc : nullptr, NP, false, FR, NT);
    if ((g.nx + 1) / 2 > NHM || nhe % NP != 0 || (TY + 2) * (nhe / NP) > NT || lds > 160 * 1024 ||
        (double)g.n * sizeof(T) >= 4294967296.0 || (PRO && (double)gc.n * sizeof(T) >= 4294967296.0))
        return 1;
    auto no_scratch = [](const void* k) {
        hipFuncAttributes a;
        return hipFuncGetAttributes(&a, k) == hipSuccess && a.localSizeBytes == 0;
    };
    const auto kwr = jsv::k_rbsweep3p<T, TY, NT, NHM, PRO, NP, 1, true, false, FR>;
    const auto krt = jsv::k_rbsweep3p<T, TY, NT, NHM, PRO, NP, 1, false, false, FR>;
    static const bool wr_ok = no_scratch((const void*)kwr), rt_ok = no_scratch((const void*)krt);
    const bool wr = (nhe / NP) % 64 == 0 && wr_ok;
    if (!wr && !rt_ok) return 1;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wr ? kwr : krt, NT, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int ytiles = (g.my - 1 + TY - 1) / TY;
    int zc = P->n_cu * per_cu / ytiles;
    if (zc < 1) zc = 1;
    if (zc > g.mz - 1) zc = g.mz - 1;
    const int zchunk = P->zchunk > 0 ? P->zchunk : (g.mz - 1 + zc - 1) / zc;
    const int zchunks = (g.mz - 1 + zchunk - 1) / zchunk;
    // zig-zag: a sweep without the prolongation reverses the level's previous
    // pair sweep; with it, only with chunks of alternating direction
    // (zig-zag 2), and otherwise the march is upward
    unsigned& last = P->zz_dn[sizeof(T) == 8 ? 0 : 1];
    const unsigned bit = 1u << l;
    const bool rev = PRO ? P->zigzag == 2 : P->zigzag != 0;
    const int dn = (rev && (last & bit) == 0u) ? 1 : 0;
    last = dn ? (last | bit) : (last & ~bit);
    const int alt = P->zigzag == 2 ? 2 : 0;   // chunks of alternating direction
    hipLaunchKernelGGL(wr ? kwr : krt, dim3((unsigned)(ytiles * zchunks)), dim3(NT), lds, s, H.phi[l], H.phi2[l],
                       H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], g, zchunk, ytiles, PRO ? pc : (const T*)nullptr,
                       gc, P->act_ptr(), (double*)nullptr, dn | alt, FR ? H.res[l] : (T*)nullptr);
    if (FR) {
        fr->ty = TY;
        fr->zchunk = zchunk;
        fr->ytiles = ytiles;
        fr->dn = alt ? 0 : dn;   // alternating chunks keep the upward march's chunk bounds
    }
    return 0;
}

// Rows of at most 260 points only: a wider row needs more than NT / (TY + 2)
// threads at two pairs per thread for these shapes.
template <typename T, int TY, int NP, int NT>
int launch_rbsweep3p_nt(const jsv_plan* P, Hier<T>& H, int l, const T* pc, hipStream_t s, FrGeo* fr)
{
    if ((H.g[l].nx + 1) / 2 > 130) return 1;
    if (fr != nullptr) return launch_rbsweep3p_nt_v<T, TY, 130, NP, NT, false, true>(P, H, l, pc, s, fr);
    if (pc != nullptr) return launch_rbsweep3p_nt_v<T, TY, 130, NP, NT, true, false>(P, H, l, pc, s, fr);
    return launch_rbsweep3p_nt_v<T, TY, 130, NP, NT, false, false>(P, H, l, pc, s, fr);
}

// Pairs per thread, rows and z-chunk of the pair sweep.  Two pairs per thread
// in FP32 (fewer halo rows per useful row and twice the work between
// barriers) where the registers allow it, one otherwise.  Untuned, every fitting row
// count is scored by the fraction of workgroup slots one wave of tiles fills
// times the share of pair rows that are not halo, TY / (TY + 2); the z-chunk
// then gives each slot one tile.
template <typename T>
int launch_rbsweep3p(const jsv_plan* P, Hier<T>& H, int l, const T* pc, hipStream_t s, double* res = nullptr,
                     long* nres = nullptr, FrGeo* fr = nullptr)
{
    const Geo& g = H.g[l];
    const Geo* gc = pc != nullptr ? &H.g[l + 1] : nullptr;
    // default: two pairs in FP32 sweeps of levels >= 256 wide (in FP64 the
    // two-pair kernel spills registers; on 128^3 one pair measured faster);
    // tuning forces
    int np = P->sweep_pairs > 0 ? P->sweep_pairs : ((sizeof(T) == 4 && g.mx >= 256) ? 2 : 1);
    if (np == 2 && !pair_sweep_rows_fit(g, 2, sizeof(T), gc, 2)) np = 1;
    // FP32 two-pair sweeps in 10-row tiles of 768 threads: the one storing
    // the first-half residuals always, the plain one by default (101 against
    // 106 us at 256^3), the one with the prolongation when tuned so (129
    // against 122 us); 1024 threads otherwise
    if constexpr (sizeof(T) == 4) {
        if (np == 2 && res == nullptr && (P->sweep_rows == 0 || P->sweep_rows == 10)) {
            const bool frv = fr != nullptr && pc == nullptr && H.res[l] != nullptr && P->fuse_resid;
            const bool plain768 = pc == nullptr && P->sweep_threads == 0;
            if (frv || plain768 || P->sweep_threads == 768) {
                if (launch_rbsweep3p_nt<T, 10, 2, 768>(P, H, l, pc, s, frv ? fr : nullptr) == 0) return 0;
            }
        }
    }
    int rows = P->sweep_rows;
    if (rows == 8 || rows == 16) rows = 6;
    const bool tuned = rows > 0 && pair_sweep_rows_fit(g, rows, sizeof(T), gc, np);
    int zchunks = 0;
    if (!tuned) {
        rows = 0;
        double best = -1.0;
        for (int ty : {14, 6, 5, 4, 2}) {
            if (!pair_sweep_rows_fit(g, ty, sizeof(T), gc, np)) continue;
            const int slots = P->n_cu * (np == 2 ? pair_sweep_per_cu<T, 2>(g, ty) : pair_sweep_per_cu<T, 1>(g, ty));
            const int ytiles = (g.my - 1 + ty - 1) / ty;
            int zc = slots / ytiles;
            if (zc < 1) zc = 1;
            if (zc > g.mz - 1) zc = g.mz - 1;
            const int used = ytiles * zc;
            const double waves = std::ceil((double)used / slots);
            const double score = ((double)used / (waves * slots)) * ty / (ty + 2.0);
            if (score > best + 1e-9) {
                best = score;
                rows = ty;
                zchunks = zc;
            }
        }
    }
    if (rows == 0) return fail("pair sweep does not fit level %d", l);
    int zchunk = P->zchunk;
    if (zchunk <= 0) {
        if (zchunks <= 0) {
            const int ytiles = (g.my - 1 + rows - 1) / rows;
            zchunks = P->n_cu * (np == 2 ? pair_sweep_per_cu<T, 2>(g, rows) : pair_sweep_per_cu<T, 1>(g, rows)) /
                      ytiles;
            if (zchunks < 1) zchunks = 1;
        }
        zchunk = (g.mz - 1 + zchunks - 1) / zchunks;
    }
    if (np == 2) return launch_rbsweep3p_rows<T, 2>(P, H, l, rows, zchunk, pc, s, res, nres, fr);
    return launch_rbsweep3p_rows<T, 1>(P, H, l, rows, zchunk, pc, s, res, nres, fr);
}

// tile rows x threads per workgroup of the fused sweep (tunable; default by precision)
template <typename T>
int launch_rbsweep3(const jsv_plan* P, Hier<T>& H, int l, hipStream_t s, const T* pc = nullptr,
                    double* res = nullptr, long* nres = nullptr, FrGeo* fr = nullptr)
{
    if (nres != nullptr) *nres = 0;
    if (fr != nullptr) fr->ty = 0;
    if (pc != nullptr) return launch_rbsweep3p<T>(P, H, l, pc, s);
    if (P->sweep_variant == 2 && pair_sweep_fits(H.g[l], sizeof(T)))
        return launch_rbsweep3p<T>(P, H, l, nullptr, s, res, nres, fr);
    int rows = P->sweep_rows > 0 ? P->sweep_rows : (sizeof(T) == 4 ? 16 : 4);
    const int nt = P->sweep_threads > 0 ? P->sweep_threads : 1024;
    if (nt == 256) {
        if (rows == 4) return launch_rbsweep3_t<T, 4, 256>(P, H, l, s);
        if (rows == 6) return launch_rbsweep3_t<T, 6, 256>(P, H, l, s);
        if (rows == 8) return launch_rbsweep3_t<T, 8, 256>(P, H, l, s);
        return launch_rbsweep3_t<T, 16, 256>(P, H, l, s);
    }
    if (rows == 4) return launch_rbsweep3_t<T, 4, 1024>(P, H, l, s);
    if (rows == 6) return launch_rbsweep3_t<T, 6, 1024>(P, H, l, s);
    if (rows == 8) return launch_rbsweep3_t<T, 8, 1024>(P, H, l, s);
    return launch_rbsweep3_t<T, 16, 1024>(P, H, l, s);
}

// The 3D post-smoothing sweep of level l can apply the prolongation from
// level l + 1 itself (pair sweep with the coarse-plane ring).
template <typename T>
bool prolong_fusable(const jsv_plan* P, const Hier<T>& H, int l)
{
    if (P->dim != 3 || P->fuse_prolong == 0 || (sizeof(T) == 8 && P->fuse_prolong < 2) || l + 1 >= P->levels)
        return false;
    const Geo& g = H.g[l];
    if (!(P->fused_min_m > 0 && g.mz >= P->fused_min_m && H.phi2[l] != nullptr)) return false;
    return P->sweep_variant == 2 && pair_sweep_fits(g, sizeof(T), &H.g[l + 1]);
}

// Level l smooths with one launch per RBGS half in the plain order (2D
// without D_xy; 3D below the fused-sweep size without the slab schedule).
template <typename T>
bool halves_path(const jsv_plan* P, const Hier<T>& H, int l)
{
    if (P->dim == 2) return !P->offdiag;
    const Geo& g = H.g[l];
    if (P->fused_min_m > 0 && g.mz >= P->fused_min_m && H.phi2[l] != nullptr) return false;
    const int planes = g.mz - 1;
    const long plane_bytes = (long)g.sz * sizeof(T) * 6;
    return !(P->slab_bytes > 0 && plane_bytes * planes > P->slab_bytes * 2);
}

// pc: coarse correction to prolongate into level l first (3D fused pair sweep,
// see prolong_fusable; or, with fuse_halves, the first half of the halves path).
// red_done: the first half already ran inside the restriction that produced
// this level (fuse_halves; halves path only), so only the second half runs.
// res / nres: see launch_rbsweep3p_t (fused out-of-place sweeps only; *nres
// is 0 where the residual was not summed)
// fr: the sweep may also store the first-half residuals of the swept phi for
// the restriction that follows (*fr says which ones; ty = 0: none)
template <typename T>
int smooth(const jsv_plan* P, Hier<T>& H, int l, hipStream_t s, const T* pc = nullptr, bool red_done = false,
           double* res = nullptr, long* nres = nullptr, FrGeo* fr = nullptr)
{
    const Geo& g = H.g[l];
    if (nres != nullptr) *nres = 0;
    if (fr != nullptr) fr->ty = 0;
    if (P->dim == 2) {
        if (P->offdiag) {
            // the reference's sub-colour order (0,0),(1,1),(0,1),(1,0)
            static const int order[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
            for (int k = 0; k < 4; ++k) {
                const int lx = order[k][0], ly = order[k][1];
                dim3 gr((unsigned)((g.mx / 2 + jsv::BX - 1) / jsv::BX),
                        (unsigned)((g.my / 2 + jsv::BY - 1) / jsv::BY), 1);
                gr.z = (unsigned)g.nz;
                hipLaunchKernelGGL(jsv::k_rb2_diag<T>, gr, dim3(jsv::BX, jsv::BY), 0, s,
                                   H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cc[l], H.drr[l],
                                   H.dlr[l], g, lx, ly, P->act_ptr());
            }
        } else {
            dim3 gr = grid_points(g, 2, g.mx / 2);
            if (pc != nullptr)
                hipLaunchKernelGGL(jsv::k_rb2_pro<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], H.rhs[l],
                                   H.cx[l], H.cy[l], H.cc[l], g, 0, P->act_ptr(), pc, H.g[l + 1]);
            else if (!red_done)
                hipLaunchKernelGGL(jsv::k_rb2<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l],
                                   H.rhs[l], H.cx[l], H.cy[l], H.cc[l], g, 0, P->act_ptr());
            hipLaunchKernelGGL(jsv::k_rb2<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l],
                               H.rhs[l], H.cx[l], H.cy[l], H.cc[l], g, 1, P->act_ptr());
        }
    } else if (P->fused_min_m > 0 && g.mz >= P->fused_min_m && H.phi2[l] != nullptr) {
        // fused out-of-place sweep, then swap the level's phi buffers
        RETCHK(launch_rbsweep3<T>(P, H, l, s, pc, res, nres, fr));
        std::swap(H.phi[l], H.phi2[l]);
    } else {
        // Slab schedule: red on z-slab k, then black on slab k-1, so the black
        // half re-reads a slab the red half fetched two launches earlier (still
        // in the Infinity Cache) instead of streaming the level twice from HBM.
        // Each point sees exactly the values of the plain red-then-black order:
        // red on slab k reads black planes not yet updated, black on slab k-1
        // reads red planes that are all updated.
        const int planes = g.mz - 1;               // unknown planes z = 2..mz
        const long plane_bytes = (long)g.sz * sizeof(T) * 6;
        int slab = planes;
        if (P->slab_bytes > 0 && plane_bytes * planes > P->slab_bytes * 2) {
            slab = (int)(P->slab_bytes / plane_bytes);
            if (slab < 2) slab = 2;
        }
        dim3 gr = grid_points(g, 3, g.mx / 2);
        if (slab == planes && (pc != nullptr || red_done)) {
            // halves path with a fused first half
            if (pc != nullptr)
                hipLaunchKernelGGL(jsv::k_rb3_pro<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], H.rhs[l],
                                   H.cx[l], H.cy[l], H.cz[l], H.cc[l], g, 1, pc, H.g[l + 1], P->act_ptr());
            hipLaunchKernelGGL(jsv::k_rb3<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], H.rhs[l], H.cx[l],
                               H.cy[l], H.cz[l], H.cc[l], g, 0, 2, P->act_ptr());
            return 0;
        }
        auto launch = [&](int par, int za, int zb) {
            dim3 gs(gr.x, gr.y, (unsigned)(zb - za));
            hipLaunchKernelGGL(jsv::k_rb3<T>, gs, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], H.rhs[l],
                               H.cx[l], H.cy[l], H.cz[l], H.cc[l], g, par, za, P->act_ptr());
        };
        const int zend = g.mz + 1;
        int prev_a = -1, prev_b = -1;
        for (int za = 2; za < zend; za += slab) {
            const int zb = za + slab < zend ? za + slab : zend;
            launch(1, za, zb);
            if (prev_a >= 0) launch(0, prev_a, prev_b);
            prev_a = za;
            prev_b = zb;
        }
        launch(0, prev_a, prev_b);
    }
    return 0;
}

// zphi: the coarse phi to zero in the same kernels (nullptr: left alone);
// red_first: also run the coarse level's first RBGS half (zphi must be set);
// fr: the sweep before stored the first-half residuals of its tiles' inner
// rows and planes (fr->ty > 0), so only the edge rows and planes remain
template <typename T>
int restrict_(const jsv_plan* P, Hier<T>& H, int l, hipStream_t s, T* zphi = nullptr, bool red_first = false,
              const FrGeo* fr = nullptr)
{
    const Geo& gf = H.g[l];
    const Geo& gc = H.g[l + 1];
    jsv::RedZ<T> rz{nullptr, nullptr, nullptr, nullptr};
    if (red_first && zphi != nullptr)
        rz = {H.cx[l + 1], H.cy[l + 1], P->dim == 3 ? H.cz[l + 1] : nullptr, H.cc[l + 1]};
    dim3 gr = grid_points(gc, P->dim, gc.mx - 1);
    if (H.res[l] != nullptr) {
        // two passes: residual at the first-half points, then the weighting
        dim3 gf_half = grid_points(gf, P->dim, gf.mx / 2);
        if (P->dim == 2) {
            hipLaunchKernelGGL(jsv::k_resid_first2<T>, gf_half, dim3(jsv::BX, jsv::BY), 0, s, H.res[l],
                               H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cc[l],
                               P->offdiag ? H.drr[l] : nullptr, P->offdiag ? H.dlr[l] : nullptr, gf,
                               P->act_ptr());
            hipLaunchKernelGGL(jsv::k_restrict_gather2<T>, gr, dim3(jsv::BX, jsv::BY), 0, s,
                               H.rhs[l + 1], H.res[l], gf, gc, P->act_ptr(), zphi, rz);
        } else {
            if (fr != nullptr && fr->ty > 0) {
                const int zchunks = (gf.mz - 1 + fr->zchunk - 1) / fr->zchunk;
                const long nb = (long)(fr->ytiles + 1) * (gf.mz - 1) + (long)(zchunks + 1) * (gf.my - 1);
                hipLaunchKernelGGL(jsv::k_resid_edges3<T>, dim3((unsigned)nb), dim3(jsv::EDGE_THREADS), 0, s,
                                   H.res[l], H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], gf, fr->ty,
                                   fr->zchunk, fr->ytiles, fr->dn, P->act_ptr());
            } else if (gf.n < (1L << 24)) {
                hipLaunchKernelGGL(jsv::k_resid_first3<T>, gf_half, dim3(jsv::BX, jsv::BY), 0, s, H.res[l],
                                   H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], gf, 1, P->act_ptr());
            } else {
                constexpr int RB = 16;
                const dim3 gt(gf_half.x, (unsigned)((gf.my - 1 + RB - 1) / RB), gf_half.z);
                hipLaunchKernelGGL((jsv::k_resid_first3<T, RB>), gt, dim3(jsv::BX, RB), 0, s, H.res[l],
                                   H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], gf, 0, P->act_ptr());
            }
            // the gather indexes the compact residuals in 32 bits
            if ((long)(gf.mz - 1) * (gf.my - 1) * (gf.mx / 2) >= (1L << 32))
                return fail("level %d too large for the two-pass restriction", l);
            hipLaunchKernelGGL(jsv::k_restrict_gather3<T>, gr, dim3(jsv::BX, jsv::BY), 0, s,
                               H.rhs[l + 1], H.res[l], gf, gc, zphi, rz, P->act_ptr());
        }
        return 0;
    }
    if (P->dim == 2)
        hipLaunchKernelGGL(jsv::k_restrict2<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.rhs[l + 1],
                           H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cc[l],
                           P->offdiag ? H.drr[l] : nullptr, P->offdiag ? H.dlr[l] : nullptr, gf, gc,
                           P->act_ptr(), zphi, rz);
    else
        hipLaunchKernelGGL(jsv::k_restrict3<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.rhs[l + 1],
                           H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], gf, gc, zphi, rz,
                           P->act_ptr());
    return 0;
}

template <typename T>
int prolong(const jsv_plan* P, Hier<T>& H, int l, int nu2, hipStream_t s)
{
    const Geo& gf = H.g[l];
    const Geo& gc = H.g[l + 1];
    dim3 gr = grid_points(gf, P->dim, gf.mx - 1);
    if (P->dim == 2) {
        const int full = (nu2 <= 0 || P->offdiag) ? 1 : 0;
        hipLaunchKernelGGL(jsv::k_prolong2<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l],
                           H.phi[l + 1], gf, gc, full, P->act_ptr());
    } else {
        hipLaunchKernelGGL(jsv::k_prolong3<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l],
                           H.phi[l + 1], gf, gc, nu2 <= 0 ? 1 : 0, P->act_ptr());
    }
    return 0;
}

template <typename T>
int zero(T* a, long n, hipStream_t s)
{
    hipLaunchKernelGGL(jsv::k_fill<T>, grid_flat(n), dim3(256), 0, s, a, n, (T)0);
    return 0;
}

// ---------------------------------------------------------------------------
// coarse-level program (k_coarse_prog): the sub-cycle below level `start - 1`
// runs as one launch of one workgroup per problem
// ---------------------------------------------------------------------------
template <typename T>
bool uses_fused(const jsv_plan* P, const Hier<T>& H, int l)
{
    return P->dim == 3 && P->fused_min_m > 0 && H.g[l].mz >= P->fused_min_m && H.phi2[l] != nullptr;
}

// first level (>= 1) handled by the coarse program, or -1
template <typename T>
int coarse_start(const jsv_plan* P, const Hier<T>& H)
{
    const int cm = P->coarse_m[sizeof(T) == 4 ? 1 : 0];
    if (cm <= 0) return -1;
    for (int l = 1; l <= P->levels - 2; ++l) {
        const Geo& g = H.g[l];
        int mm = g.mx > g.my ? g.mx : g.my;
        if (P->dim == 3 && g.mz > mm) mm = g.mz;
        if (mm > cm) continue;
        for (int k = l; k < P->levels; ++k)
            if (uses_fused(P, H, k)) return -1;
        return l;
    }
    return -1;
}

// the op list of mg_cycle(l, fcyc, skip_pre = false), in mg_cycle's order.
// With fuse_halves (2D without D_xy) the launch levels' two fusions apply here
// too: the restriction into level l + 1 also runs the first half of the first
// sweep of its first visit (phi = 0 there), and the prolongation runs inside
// the first post-smoothing half (nu2 > 0).  red_done: the restriction into
// level l already ran the first half of this visit's first sweep.  In 3D the
// fused post-smoothing half (six corrected neighbours per point, two to eight
// coarse values each) made the program slower (FP32 34 -> 39 us per launch)
// and the fused restriction gained nothing, so 3D programs keep separate ops.
bool op_empty(const jsv_plan* P, const std::vector<Geo>& G, const jsv::CoarseOp& op);

void build_ops(const jsv_plan* P, const std::vector<Geo>& G, const jsv_config& c, int l, bool fcyc,
               std::vector<jsv::CoarseOp>& ops, bool red_done = false)
{
    const bool fuse = P->fuse_halves && P->dim == 2 && !P->offdiag;
    const int first_par = P->dim == 3 ? 1 : 0;    // parity of a sweep's first half
    auto smooth_ops = [&](int lv, bool skip_first) {
        if (P->dim == 2 && P->offdiag) {
            static const int order[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
            for (int k = 0; k < 4; ++k) ops.push_back({jsv::CO_RB_DIAG, lv, order[k][0], order[k][1]});
            return;
        }
        if (!skip_first) ops.push_back({jsv::CO_RB, lv, first_par, 0});
        ops.push_back({jsv::CO_RB, lv, 1 - first_par, 0});
    };
    if (l == P->levels - 2) {
        for (int i = 0; i < c.nu1 + c.nu2; ++i) smooth_ops(l, red_done && i == 0);
        return;
    }
    for (int i = 0; i < c.nu1; ++i) smooth_ops(l, red_done && i == 0);
    const int first_sweeps = (l + 1 == P->levels - 2) ? c.nu1 + c.nu2 : c.nu1;
    const bool red = fuse && first_sweeps >= 1;
    ops.push_back({jsv::CO_RESTRICT, l, red ? 1 : 0, 1});      // also phi of l + 1 := 0
    bool first = red;
    if (fcyc)
        for (int i = 0; i < c.gamma; ++i) {
            build_ops(P, G, c, l + 1, true, ops, first);
            first = false;
        }
    for (int i = 0; i < c.gamma; ++i) {
        build_ops(P, G, c, l + 1, false, ops, first);
        first = false;
    }
    const int full = (c.nu2 <= 0 || (P->dim == 2 && P->offdiag)) ? 1 : 0;
    if (fuse && c.nu2 >= 1 && !op_empty(P, G, {jsv::CO_RB, l, first_par, 0})) {
        ops.push_back({jsv::CO_PRO_RB, l, first_par, 0});
        ops.push_back({jsv::CO_RB, l, 1 - first_par, 0});
        for (int i = 1; i < c.nu2; ++i) smooth_ops(l, false);
        return;
    }
    ops.push_back({jsv::CO_PROLONG, l, full, 0});
    for (int i = 0; i < c.nu2; ++i) smooth_ops(l, false);
}

template <typename T>
const jsv_plan::CoarseProg* find_prog(const jsv_plan* P, const jsv_config& c, bool fcyc, int start)
{
    const int prec = sizeof(T) == 4 ? JSV_MIXED : JSV_FP64;
    const int pi = sizeof(T) == 4 ? 1 : 0;
    for (const auto& pr : P->progs)
        if (pr.prec == prec && pr.gamma == c.gamma && pr.nu1 == c.nu1 && pr.nu2 == c.nu2 &&
            pr.fcyc == (fcyc ? 1 : 0) && pr.start == start && pr.offdiag == (P->offdiag ? 1 : 0) &&
            pr.lds_from == P->clds_from[pi] && pr.solo == P->coarse_solo && pr.reg == P->coarse_reg &&
            pr.fuse == P->fuse_halves)
            return &pr;
    return nullptr;
}

// A smoothing op with no point of its colour on the level (restriction and
// prolongation always have points).
bool op_empty(const jsv_plan* P, const std::vector<Geo>& G, const jsv::CoarseOp& op)
{
    const Geo& g = G[op.level];
    if (op.kind == jsv::CO_RB_DIAG) {
        // x = (a ? 2 : 3) + 2i <= mx, y = (b ? 2 : 3) + 2j <= my
        const int nx = op.a ? g.mx / 2 : (g.mx - 1) / 2, ny = op.b ? g.my / 2 : (g.my - 1) / 2;
        return nx <= 0 || ny <= 0;
    }
    if (op.kind != jsv::CO_RB) return false;
    // x = 2 + ((a + y [+ z]) & 1) + 2i <= mx for some y (and z) in [2, m]
    for (int z = 2; z <= (P->dim == 3 ? g.mz : 2); ++z)
        for (int y = 2; y <= g.my; ++y)
            if (2 + ((op.a + y + (P->dim == 3 ? z : 0)) & 1) <= g.mx) return false;
    return true;
}

// Loop count of one op in k_coarse_prog, for the one-wave runs: power-of-two
// extents (the resolutions); a zeroing op covers the whole array (compact on
// staged levels, ln elements).
long op_points(const jsv_plan* P, const std::vector<Geo>& G, const std::vector<int>& ln, bool staged,
               const jsv::CoarseOp& op)
{
    const Geo& g = G[op.level];
    const long nz = P->dim == 3 ? g.mz : 1;
    const long full = staged ? ln[op.level] : (P->dim == 3 ? g.n : g.sz);
    switch (op.kind) {
    case jsv::CO_RB:
    case jsv::CO_RB_PAIR:
    case jsv::CO_PRO_RB: return (long)(g.mx / 2) * g.my * nz;
    case jsv::CO_RB_DIAG: return (long)(g.mx / 2) * (g.my / 2);
    case jsv::CO_ZERO: return full;
    case jsv::CO_PROLONG: return (long)g.mx * g.my * nz;
    default: {
        const Geo& gc = G[op.level + 1];
        const long inner = (long)gc.mx * gc.my * (P->dim == 3 ? gc.mz : 1);
        const long zc = op.b ? (staged ? ln[op.level + 1] : (P->dim == 3 ? gc.n : gc.sz)) : 0;
        return inner > zc ? inner : zc;
    }
    }
}

// Uploads the level table and the two op lists a cycle of configuration c
// needs.  Synchronous copies: never called while a stream is being captured.
template <typename T>
int prepare_coarse(jsv_plan* P, Hier<T>& H, const jsv_config& c, hipStream_t s)
{
    const int start = coarse_start(P, H);
    if (start < 0) return 0;
    const int pi = sizeof(T) == 4 ? 1 : 0;
    // only levels >= start are read by the program (levels above may swap
    // their phi buffers every sweep)
    // D_xy arrays only while the operator has them: they stay allocated (with
    // stale values) after the plan switches back to a 5-point operator
    auto drr = [&](int l) -> T* { return P->offdiag ? H.drr[l] : nullptr; };
    auto dlr = [&](int l) -> T* { return P->offdiag ? H.dlr[l] : nullptr; };
    // LDS staging: the coarsest levels (not level L-1, which no cycle
    // visits) from the bottom up while they fit.  Each staged level is one
    // block of its arrays, each compact over indices 1..m+1 per axis; the
    // whole region arrives as one image (coefficients, zero phi / rhs).
    int lds_from = P->levels - 1;
    const int lds_to = P->levels - 1;
    auto has = [&](int j) { return j <= 3 || j == 5 || (j == 4 && P->dim == 3) || (j >= 6 && P->offdiag); };
    std::vector<int> lsy(P->levels), lsz(P->levels), ln(P->levels);
    for (int l = 0; l < P->levels; ++l) {
        const Geo& g = H.g[l];
        lsy[l] = g.mx + 1;
        lsz[l] = lsy[l] * (g.my + 1);
        ln[l] = P->dim == 3 ? lsz[l] * (g.mz + 1) : lsz[l];
    }
    // op lists of both programs (V/W and F entry) first: their LDS copy shares the budget
    std::vector<jsv::CoarseOp> ops_of[2];
    for (int fc = 0; fc < 2; ++fc) {
        std::vector<jsv::CoarseOp>& ops = ops_of[fc];
        ops.push_back({jsv::CO_ZERO, start, 0, 0, 0});
        if (fc)
            for (int i = 0; i < c.gamma; ++i) build_ops(P, H.g, c, start, true, ops);
        for (int i = 0; i < c.gamma; ++i) build_ops(P, H.g, c, start, false, ops);
    }
    const size_t op_bytes = 4 * std::max(ops_of[0].size(), ops_of[1].size());
    size_t nar = 0;
    for (int j = 0; j < 8; ++j) nar += has(j) ? 1 : 0;
    size_t used = 0;
    if (P->coarse_lds && op_bytes <= 4 * (size_t)jsv::COARSE_MAX_OPS) {
        for (int l = P->levels - 2; l >= start; --l) {
            const size_t lb = (size_t)ln[l] * nar * sizeof(T);
            if (used + lb + op_bytes > 156 * 1024) break;
            used += lb;
            lds_from = l;
        }
    }
    std::vector<std::array<int, 8>> loff(P->levels);
    for (auto& a : loff) a.fill(-1);
    // the start level's rhs goes last, after the image: the kernel copies it
    // in from the restriction's output in the same pass as the image
    const bool start_staged = start >= lds_from && start < lds_to;
    int o = 0;
    for (int l = lds_from; l < lds_to; ++l)
        for (int j = 0; j < 8; ++j)
            if (has(j) && !(start_staged && l == start && j == 1)) {
                loff[l][j] = o;
                o += ln[l];
            }
    const int img_n = (int)((o * sizeof(T) + 15) / 16 * 16 / sizeof(T));   // whole 16-byte pieces
    o = img_n;
    if (start_staged) {
        loff[start][1] = o;
        o += ln[start];
    }
    used = (size_t)o * sizeof(T);
    P->clds_from[pi] = lds_from;
    P->clds_to[pi] = lds_to;
    P->clds_bytes[pi] = used;
    P->img_n[pi] = img_n;
    std::vector<const void*> key{(const void*)(long)start, (const void*)(long)lds_from};
    for (int l = start; l < P->levels; ++l)
        for (const T* q : {H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], drr(l), dlr(l)})
            key.push_back(q);
    if (P->lv_host[pi] != key) {
        std::vector<jsv::LvPtr<T>> lv(P->levels);
        for (int l = start; l < P->levels; ++l) {
            lv[l] = {H.phi[l], H.rhs[l], H.cx[l], H.cy[l], H.cz[l], H.cc[l], drr(l), dlr(l), H.g[l], {},
                     lsy[l], lsz[l], ln[l]};
            for (int j = 0; j < 8; ++j) lv[l].lds[j] = loff[l][j];
        }
        const size_t nb = sizeof(jsv::LvPtr<T>) * lv.size();
        if (P->lv_dev[pi] == nullptr) RETCHK(dalloc(P, &P->lv_dev[pi], nb));
        HIPCHK(hipStreamSynchronize(s));   // a queued cycle may still read the table
        HIPCHK(hipMemcpy(P->lv_dev[pi], lv.data(), nb, hipMemcpyHostToDevice));
        P->lv_host[pi] = key;
    }
    // the image follows the coefficients: rebuilt when the operator or the layout changed
    if (img_n > 0) {
        std::vector<const void*> ikey = key;
        ikey.push_back((const void*)P->op_version);
        if (P->img_key[pi] != ikey) {
            const size_t nb = sizeof(T) * (size_t)img_n * (size_t)P->batch;
            if (P->img_cap[pi] < nb) {
                // captured cycles hold the old image's address: drop them with it
                HIPCHK(hipStreamSynchronize(s));
                drop_graphs(P);
                RETCHK(dfree(P, P->img_dev[pi], P->img_cap[pi]));
                P->img_dev[pi] = nullptr;
                RETCHK(dalloc(P, &P->img_dev[pi], nb));
                P->img_cap[pi] = nb;
            }
            T* img = (T*)P->img_dev[pi];
            RETCHK(zero(img, (long)img_n * P->batch, s));      // phi and rhs slots
            for (int l = lds_from; l < lds_to; ++l) {
                const T* src[8] = {nullptr, nullptr, H.cx[l], H.cy[l], H.cz[l], H.cc[l], drr(l), dlr(l)};
                const long sbs = P->dim == 2 ? H.g[l].sz : 0;
                for (int j = 2; j < 8; ++j)
                    if (has(j))
                        hipLaunchKernelGGL(jsv::k_stage_image<T>,
                                           dim3((unsigned)((ln[l] + 255) / 256), (unsigned)P->batch), dim3(256), 0,
                                           s, img, (long)img_n, loff[l][j], src[j], sbs, H.g[l], lsy[l], lsz[l],
                                           ln[l], P->dim);
            }
            LAUNCHCHK();
            P->img_key[pi] = ikey;
        }
    }
    for (int fc = 0; fc < 2; ++fc) {
        if (find_prog<T>(P, c, fc != 0, start) != nullptr) continue;
        std::vector<jsv::CoarseOp>& ops = ops_of[fc];
        if (ops.size() > (size_t)jsv::COARSE_MAX_OPS) {
            P->progs.push_back({sizeof(T) == 4 ? JSV_MIXED : JSV_FP64, c.gamma, c.nu1, c.nu2, fc, start,
                                P->offdiag ? 1 : 0, lds_from, P->coarse_solo, P->coarse_reg, P->fuse_halves,
                                nullptr, (int)ops.size()});
            continue;
        }
        // Ops that change nothing go: smoothing halves without points, and on
        // a level whose only unknown neighbours boundary nodes every
        // smoothing op after the first of a visit (the unknown's update reads
        // only rhs and coefficients, so it recomputes the same bits; a
        // restriction that runs the level's first half counts as that
        // first).  The first op zeroes a staged start level's phi, which the
        // image already holds as zeros.
        auto single = [&](int l) {
            const Geo& g = H.g[l];
            return g.mx == 2 && g.my == 2 && (P->dim == 2 || g.mz == 2);
        };
        {
            std::vector<jsv::CoarseOp> kept;
            bool done_single = false;   // the single unknown was updated in this visit
            for (const jsv::CoarseOp& op : ops) {
                const bool smooth = op.kind == jsv::CO_RB || op.kind == jsv::CO_RB_DIAG;
                if (smooth && op_empty(P, H.g, op)) continue;
                if (smooth && single(op.level)) {
                    if (done_single) continue;
                    done_single = true;
                } else {
                    // the single unknown (2, 2[, 2]) has even parity: the first
                    // half in 2D, which a restriction with a = 1 runs; in 3D it
                    // is the second half's
                    done_single = op.kind == jsv::CO_RESTRICT && op.a == 1 && single(op.level + 1) && P->dim == 2;
                }
                if (op.kind == jsv::CO_ZERO && op.level == start && start >= lds_from && start < lds_to &&
                    kept.empty())
                    continue;
                kept.push_back(op);
            }
            ops.swap(kept);
        }
        // runs of small ops on staged levels go to one wave
        auto staged = [&](int l) { return l >= lds_from && l < lds_to; };
        auto solo = [&](const jsv::CoarseOp& op) {
            const bool two = op.kind == jsv::CO_RESTRICT || op.kind == jsv::CO_PROLONG || op.kind == jsv::CO_PRO_RB;
            return P->coarse_solo > 0 && staged(op.level) && (!two || staged(op.level + 1)) &&
                   op_points(P, H.g, ln, true, op) <= P->coarse_solo;
        };
        {
            // in a wave's run, the two halves of one sweep share one op
            std::vector<jsv::CoarseOp> merged;
            for (size_t i = 0; i < ops.size(); ++i) {
                const jsv::CoarseOp& a = ops[i];
                if (i + 1 < ops.size() && a.kind == jsv::CO_RB && ops[i + 1].kind == jsv::CO_RB &&
                    ops[i + 1].level == a.level && ops[i + 1].a == 1 - a.a && solo(a) && solo(ops[i + 1])) {
                    merged.push_back({jsv::CO_RB_PAIR, a.level, a.a, 0, 0});
                    ++i;
                } else if (i + 1 < ops.size() && a.kind == jsv::CO_PRO_RB && ops[i + 1].kind == jsv::CO_RB &&
                           ops[i + 1].level == a.level && ops[i + 1].a == 1 - a.a && solo(a) && solo(ops[i + 1])) {
                    merged.push_back({jsv::CO_PRO_RB, a.level, a.a, 1, 0});
                    ++i;
                } else {
                    merged.push_back(a);
                }
            }
            ops.swap(merged);
        }
        // a run keeps its levels' points in registers when it touches at most
        // three consecutive 2D levels of at most 64 points (padded), or
        // (coarse_reg 2) a level of at most 256 points with even extents and
        // up to three such levels below it, each half the one above per axis;
        // and no D_xy sub-colours
        auto wave_pts = [&](int l) { return (long)H.g[l].mx * H.g[l].my; };
        auto reg_ok = [&](size_t i, size_t k) {
            if (!P->coarse_reg || P->dim != 2 || P->offdiag) return false;
            int lmin = 1 << 30, lmax = -1;
            for (size_t j = i; j < k; ++j) {
                const jsv::CoarseOp& q = ops[j];
                if (q.kind == jsv::CO_RB_DIAG) return false;
                const int hi = (q.kind == jsv::CO_RESTRICT || q.kind == jsv::CO_PROLONG || q.kind == jsv::CO_PRO_RB)
                                   ? q.level + 1
                                   : q.level;
                lmin = std::min(lmin, q.level);
                lmax = std::max(lmax, hi);
            }
            const bool four = wave_pts(lmin) > jsv::COARSE_WAVE;
            if (four) {
                const Geo& g = H.g[lmin];
                if (P->coarse_reg < 2 || wave_pts(lmin) > 4 * jsv::COARSE_WAVE || (g.mx & 1) || (g.my & 1) ||
                    lmax - lmin > 3)
                    return false;
                for (int l = lmin + 1; l <= lmax; ++l)
                    if (2 * H.g[l].mx != H.g[l - 1].mx || 2 * H.g[l].my != H.g[l - 1].my) return false;
            } else if (lmax - lmin > 2) {
                return false;
            }
            for (int l = lmin; l <= lmax; ++l)
                if (!staged(l) || ((l > lmin || !four) && wave_pts(l) > jsv::COARSE_WAVE)) return false;
            return true;
        };
        // an op of such a run that one wave would not run alone otherwise
        auto solo4 = [&](const jsv::CoarseOp& op) {
            const bool two = op.kind == jsv::CO_RESTRICT || op.kind == jsv::CO_PROLONG || op.kind == jsv::CO_PRO_RB;
            return P->coarse_solo > 0 && P->coarse_reg >= 2 && P->dim == 2 && !P->offdiag &&
                   op.kind != jsv::CO_RB_DIAG &&
                   staged(op.level) && (!two || staged(op.level + 1)) &&
                   wave_pts(op.level) <= 4 * jsv::COARSE_WAVE;
        };
        for (size_t i = 0; i < ops.size();) {
            size_t k = i;
            while (k < ops.size() && solo(ops[k])) ++k;
            // a longer run that includes ops on a level of up to 256 points
            size_t k4 = i;
            while (k4 < ops.size() && (solo(ops[k4]) || solo4(ops[k4]))) ++k4;
            if (k4 > k && reg_ok(i, k4)) k = k4;
            if (k > i) {
                ops[i].run = (int)(k - i);
                ops[i].reg = reg_ok(i, k) ? 1 : 0;
                i = k;
            } else {
                ++i;
            }
        }
        std::vector<unsigned> words(ops.size());
        for (size_t i = 0; i < ops.size(); ++i) words[i] = jsv::coarse_pack(ops[i]);
        unsigned* d = nullptr;
        RETCHK(dalloc(P, (void**)&d, sizeof(unsigned) * words.size()));
        HIPCHK(hipMemcpy(d, words.data(), sizeof(unsigned) * words.size(), hipMemcpyHostToDevice));
        P->progs.push_back({sizeof(T) == 4 ? JSV_MIXED : JSV_FP64, c.gamma, c.nu1, c.nu2, fc, start,
                            P->offdiag ? 1 : 0, lds_from, P->coarse_solo, P->coarse_reg, P->fuse_halves, d,
                            (int)ops.size()});
    }
    return 0;
}

// The staged coefficient image holds the plan's current operator (or nothing
// is staged).  A program may only run while this holds.
template <typename T>
bool image_current(const jsv_plan* P)
{
    const int pi = sizeof(T) == 4 ? 1 : 0;
    if (P->img_n[pi] <= 0) return true;
    return !P->img_key[pi].empty() && P->img_key[pi].back() == (const void*)P->op_version;
}

template <typename T>
int mg_cycle(const jsv_plan* P, Hier<T>& H, const jsv_config& c, int l, bool fcyc, bool skip_pre,
             hipStream_t s, bool red_done = false);

// The device stop test of the relative residual (StopCtl mode 0) for the
// iterate `phi` a cycle starts from: nres partial sums of r^2 already in
// P->partials (summed inside the first sweep), or, with nres = 0, a residual
// pass over phi first.
template <typename T>
int stop_test_relres(const jsv_plan* P, Hier<T>& H, const T* phi, long nres, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        const Geo& g = H.g[0];
        if (nres == 0) {
            dim3 gr = grid_points(g, P->dim, g.mx - 1);
            hipLaunchKernelGGL((jsv::k_residual<double, double, double>), gr, dim3(jsv::BX, jsv::BY), 0, s, phi,
                               H.rhs[0], H.cx[0], H.cy[0], P->dim == 3 ? H.cz[0] : nullptr, H.cc[0],
                               P->offdiag ? H.drr[0] : nullptr, P->offdiag ? H.dlr[0] : nullptr, g, (double*)nullptr,
                               P->partials, P->dim, (const int*)P->d_act);
            nres = (long)gr.x * gr.y * gr.z;
        }
        hipLaunchKernelGGL(jsv::k_stop_check, dim3(1), dim3(1024), 0, s, P->partials, nres, 0, P->d_stop, P->d_act);
        return 0;
    }
    return fail("internal: the device stop test runs on FP64 cycles");
}

// multigrid_cycle (solver_2d.py:582-614 / solver_3d.py:588-620)
// red_done: the restriction into this level already ran the first half of
// its first sweep (fuse_halves).  test_first: level 0's first sweep is out of
// place; it sums the residual of its input where it can, and the device stop
// test runs right after it.
template <typename T>
int mg_cycle_body(const jsv_plan* P, Hier<T>& H, const jsv_config& c, int l, bool fcyc, bool skip_pre,
                  hipStream_t s, bool red_done, bool test_first)
{
    // the first sweep of the cycle, with the stop test behind it when asked
    auto first_sweep = [&]() -> int {
        const T* in = H.phi[l];
        long nres = 0;
        RETCHK(smooth(P, H, l, s, (const T*)nullptr, red_done, P->partials, &nres));
        return stop_test_relres(P, H, in, nres, s);
    };
    if (l == P->levels - 2) {
        for (int i = 0; i < c.nu1 + c.nu2; ++i) {
            if (i == 0 && test_first) RETCHK(first_sweep());
            else RETCHK(smooth(P, H, l, s, (const T*)nullptr, red_done && i == 0));
        }
        return 0;
    }
    // the last pre-smoothing sweep may store the restriction's residuals too
    FrGeo frg;
    if (!skip_pre)
        for (int i = 0; i < c.nu1; ++i) {
            if (i == 0 && test_first) RETCHK(first_sweep());
            else
                RETCHK(smooth(P, H, l, s, (const T*)nullptr, red_done && i == 0, nullptr, nullptr,
                              i == c.nu1 - 1 ? &frg : nullptr));
        }
    const int start = coarse_start(P, H);
    const jsv_plan::CoarseProg* pr = (start == l + 1) ? find_prog<T>(P, c, fcyc, start) : nullptr;
    if (pr != nullptr && pr->ops == nullptr) pr = nullptr;
    if (pr != nullptr && !image_current<T>(P)) pr = nullptr;   // staged coefficients of another operator
    // the first visit of level l + 1 starts from phi = 0 with a sweep: its first
    // half can run inside the restriction
    const int first_sweeps = (l + 1 == P->levels - 2) ? c.nu1 + c.nu2 : c.nu1;
    const bool red_first = pr == nullptr && P->fuse_halves && first_sweeps >= 1 && halves_path(P, H, l + 1);
    // phi of level l + 1 := 0 happens inside the restriction (or the program)
    RETCHK(restrict_(P, H, l, s, pr != nullptr ? (T*)nullptr : H.phi[l + 1], red_first, &frg));
    if (pr != nullptr) {
        // zero + the gamma (F: 2 gamma) sub-cycles of level l + 1 in one launch
        const int pi = sizeof(T) == 4 ? 1 : 0;
        const int lf = P->clds_from[pi], lt = P->clds_to[pi];
        const dim3 grid((unsigned)(P->dim == 2 ? P->batch : 1));
        const auto* lvp = (const jsv::LvPtr<T>*)P->lv_dev[pi];
        // LDS: the staged levels (the image, then the start level's rhs), then the op words
        const int ops_at = (int)(P->clds_bytes[pi] / sizeof(T));
        const size_t lds = P->clds_bytes[pi] + sizeof(unsigned) * (size_t)pr->nops;
        const int wb = (lf <= start && start < lt) ? start : -1;
        const T* img = (const T*)P->img_dev[pi];
        const int imn = P->img_n[pi];
        auto launch = [&](auto kern, int nt) {
            hipLaunchKernelGGL(kern, grid, dim3(nt), lds, s, lvp, pr->ops, pr->nops, ops_at, P->act_ptr(), lf, lt,
                               wb, img, imn, (long)imn);
        };
        // 2D FP64 register runs from 256-point levels need more than the 128
        // VGPRs a 1024-thread workgroup leaves a wave
        const int nt = P->coarse_threads != 0                  ? P->coarse_threads
                       : (P->dim == 2 && sizeof(T) == 8) ? 512
                                                         : jsv::COARSE_THREADS;
        if (P->dim == 2)
            nt == 256   ? launch(jsv::k_coarse_prog<T, 256, 2>, nt)
            : nt == 512 ? launch(jsv::k_coarse_prog<T, 512, 2>, nt)
                        : launch(jsv::k_coarse_prog<T, 1024, 2>, nt);
        else
            nt == 256   ? launch(jsv::k_coarse_prog<T, 256, 3>, nt)
            : nt == 512 ? launch(jsv::k_coarse_prog<T, 512, 3>, nt)
                        : launch(jsv::k_coarse_prog<T, 1024, 3>, nt);
    } else {
        bool first = red_first;
        if (fcyc)
            for (int i = 0; i < c.gamma; ++i) {
                RETCHK(mg_cycle(P, H, c, l + 1, true, false, s, first));
                first = false;
            }
        for (int i = 0; i < c.gamma; ++i) {
            RETCHK(mg_cycle(P, H, c, l + 1, false, false, s, first));
            first = false;
        }
    }
    if (c.nu2 >= 1 && (prolong_fusable(P, H, l) || (P->fuse_halves && halves_path(P, H, l)))) {
        // the first post-smoothing sweep applies the prolongation itself
        RETCHK(smooth(P, H, l, s, (const T*)H.phi[l + 1]));
        for (int i = 1; i < c.nu2; ++i) RETCHK(smooth(P, H, l, s));
        return 0;
    }
    RETCHK(prolong(P, H, l, c.nu2, s));
    for (int i = 0; i < c.nu2; ++i) RETCHK(smooth(P, H, l, s));
    return 0;
}

// One cycle from level l.  The top level of an FP64 solve whose stop rule runs
// on the device (P->stop_mode) also carries the test (k_stop_check): for the
// relative residual, of the iterate the cycle starts from -- summed inside
// the cycle's first sweep when that sweep is out of place (its input stays
// intact), else by a residual pass before the cycle; for the distance, of
// the cycle's move, from a copy of its input.
template <typename T>
int mg_cycle(const jsv_plan* P, Hier<T>& H, const jsv_config& c, int l, bool fcyc, bool skip_pre,
             hipStream_t s, bool red_done)
{
    if (sizeof(T) != 8 || l != 0 || P->stop_mode == 0)
        return mg_cycle_body(P, H, c, l, fcyc, skip_pre, s, red_done, false);
    const Geo& g = H.g[0];
    if (P->stop_mode == 2) {
        const long n = g.n;
        const unsigned nb = (unsigned)std::min<long>(std::min<long>((n + 255) / 256, 4L * P->n_cu), P->npartials);
        hipLaunchKernelGGL(jsv::k_stop_copy, dim3(nb), dim3(256), 0, s, (double*)P->phi_old, (const double*)H.phi[0],
                           n, (const jsv::StopCtl*)P->d_stop, (const int*)P->d_act);
        RETCHK(mg_cycle_body(P, H, c, 0, fcyc, skip_pre, s, red_done, false));
        hipLaunchKernelGGL(jsv::k_stop_maxdiff, dim3(nb), dim3(256), 0, s, (const double*)P->phi_old,
                           (const double*)H.phi[0], n, P->partials, (const int*)P->d_act);
        hipLaunchKernelGGL(jsv::k_stop_check, dim3(1), dim3(1024), 0, s, P->partials, (long)nb, 1, P->d_stop,
                           P->d_act);
        return 0;
    }
    const bool sweeps_first = (P->levels == 2) ? c.nu1 + c.nu2 >= 1 : (!skip_pre && c.nu1 >= 1);
    if (sweeps_first && uses_fused(P, H, 0)) return mg_cycle_body(P, H, c, 0, fcyc, skip_pre, s, red_done, true);
    RETCHK(stop_test_relres(P, H, H.phi[0], 0, s));
    return mg_cycle_body(P, H, c, 0, fcyc, skip_pre, s, red_done, false);
}

void adj_drop_graphs(const jsv_plan* P);
void adj_forget(const jsv_plan* P);

void drop_graphs(jsv_plan* P)
{
    for (auto& gr : P->graphs)
        if (gr.exec != nullptr) (void)hipGraphExecDestroy(gr.exec);   // nullptr: replayed eagerly
    P->graphs.clear();
    adj_drop_graphs(P);
}

// One top-level cycle.  The kernel sequence of a cycle depends only on the
// configuration, so it is captured once into a hipGraph (on the plan's private
// stream; capture on the legacy default stream is not allowed) and replayed on
// the caller's stream: one submission instead of ~50 launches, which keeps the
// coarse levels' tiny kernels from starving the GPU.  The cycle's kernels
// record no memset or memcpy nodes.  A configuration whose sweeps leave a
// level's phi in its second buffer (odd sweep count per visit) is not graph
// replayable and runs eagerly.
template <typename T>
int run_cycle(jsv_plan* P, Hier<T>& H, const jsv_config& c, bool skip_pre, hipStream_t s)
{
    const int prec = sizeof(T) == 4 ? JSV_MIXED : JSV_FP64;
    RETCHK(prepare_coarse(P, H, c, s));
    if (!P->use_graphs) return mg_cycle(P, H, c, 0, c.fcycle != 0, skip_pre, s);
    const int ua = P->use_act ? 1 : 0;
    for (auto& gr : P->graphs)
        if (gr.prec == prec && gr.gamma == c.gamma && gr.fcycle == (c.fcycle != 0) &&
            gr.skip_pre == (skip_pre ? 1 : 0) && gr.nu1 == c.nu1 && gr.nu2 == c.nu2 && gr.use_act == ua &&
            gr.stop_mode == P->stop_mode) {
            if (gr.exec == nullptr) return mg_cycle(P, H, c, 0, c.fcycle != 0, skip_pre, s);
            HIPCHK(hipGraphLaunch(gr.exec, s));
            return 0;
        }
    if (P->cap_stream == nullptr) HIPCHK(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
    // the capture stream must not run ahead of work already queued on s
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, s));
    HIPCHK(hipStreamWaitEvent(P->cap_stream, ev, 0));
    (void)hipEventDestroy(ev);
    // the capture runs the host side of the cycle: out-of-place sweeps swap
    // phi and phi2 of their level, so both lists are restored after it
    std::vector<T*> before = H.phi, before2 = H.phi2;
    HIPCHK(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = mg_cycle(P, H, c, 0, c.fcycle != 0, skip_pre, P->cap_stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(P->cap_stream, &graph);
    if (rc != 0) {
        if (graph) (void)hipGraphDestroy(graph);
        H.phi = before;
        H.phi2 = before2;
        return rc;
    }
    if (e != hipSuccess) {
        H.phi = before;
        H.phi2 = before2;
        return fail("graph capture failed: %s", hipGetErrorString(e));
    }
    jsv_plan::Graph gr{prec, c.gamma, c.fcycle != 0 ? 1 : 0, skip_pre ? 1 : 0, c.nu1, c.nu2, ua, P->stop_mode,
                       nullptr};
    const bool same = (H.phi == before);
    if (same) {
        e = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail("graph instantiate failed: %s", hipGetErrorString(e));
    } else {
        (void)hipGraphDestroy(graph);
        H.phi = before;                          // replay eagerly instead
        H.phi2 = before2;
    }
    P->graphs.push_back(gr);
    if (gr.exec != nullptr) {
        HIPCHK(hipGraphLaunch(gr.exec, s));
        return 0;
    }
    return mg_cycle(P, H, c, 0, c.fcycle != 0, skip_pre, s);
}

int zero_rim_level0(jsv_plan* P, double* a, hipStream_t s)
{
    const Geo& g = P->h64.g[0];
    for (int ph = 0; ph < P->dim; ++ph)
        hipLaunchKernelGGL(jsv::k_bcs, grid_exact(g.n), dim3(256), 0, s, a, g, P->dim, 1, -1,
                           nullptr, nullptr, nullptr, nullptr, ph);
    return 0;
}

// Per-problem scalar slots: slot k of problem b lives at dscal[k * batch + b].
// 3D levels are one problem; 2D levels hold `batch` problems along z.
inline int nseg(const jsv_plan* P) { return P->dim == 3 ? 1 : P->batch; }

// ||a||^2 over the unknowns of each problem of the fine level -> slot
template <typename T>
int norm2_unknowns(jsv_plan* P, const T* a, int slot, hipStream_t s)
{
    const Geo& g = P->h64.g[0];
    dim3 gr = grid_points(g, P->dim, g.mx - 1);
    hipLaunchKernelGGL(jsv::k_norm2_unknowns<T>, gr, dim3(jsv::BX, jsv::BY), 0, s, a, g,
                       P->partials, P->dim);
    const long per = (long)gr.x * gr.y * gr.z / nseg(P);
    hipLaunchKernelGGL(jsv::k_sum_partials, dim3((unsigned)nseg(P)), dim3(1024), 0, s, P->partials,
                       per, P->dscal + (long)slot * P->batch);
    return 0;
}

// FP64 residual of the fine level; optional store (double or float) and
// ||r||^2 per problem -> slot
template <typename TO>
int residual64(jsv_plan* P, const double* phi, const double* rhs, TO* rout, int slot, hipStream_t s)
{
    Hier<double>& H = P->h64;
    const Geo& g = H.g[0];
    dim3 gr = grid_points(g, P->dim, g.mx - 1);
    hipLaunchKernelGGL((jsv::k_residual<double, double, TO>), gr, dim3(jsv::BX, jsv::BY), 0, s,
                       phi, rhs, H.cx[0], H.cy[0], P->dim == 3 ? H.cz[0] : nullptr, H.cc[0],
                       P->offdiag ? H.drr[0] : nullptr, P->offdiag ? H.dlr[0] : nullptr, g, rout,
                       P->partials, P->dim, (const int*)nullptr);
    const long per = (long)gr.x * gr.y * gr.z / nseg(P);
    hipLaunchKernelGGL(jsv::k_sum_partials, dim3((unsigned)nseg(P)), dim3(1024), 0, s, P->partials,
                       per, P->dscal + (long)slot * P->batch);
    return 0;
}

// max|a-b| over each problem's whole array -> slot
int maxdiff(jsv_plan* P, const double* a, const double* b, int slot, hipStream_t s)
{
    const int ns = nseg(P);
    const long n = P->h64.g[0].n / ns;
    unsigned long long* out = reinterpret_cast<unsigned long long*>(P->dscal + (long)slot * P->batch);
    hipLaunchKernelGGL(jsv::k_fill<unsigned long long>, dim3(1), dim3(256), 0, s, out, (long)ns, 0ull);
    long nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(jsv::k_maxabsdiff<double>, dim3((unsigned)nb, (unsigned)ns), dim3(256), 0, s,
                       a, b, n, out);
    return 0;
}

// copy slots [0, nslots) of every problem to the host and wait
int read_scalars(jsv_plan* P, int nslots, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(P->hscal, P->dscal, sizeof(double) * nslots * P->batch,
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int check_cfg(const jsv_config* c)
{
    if (c == nullptr) return fail("config is NULL");
    if (c->gamma < 1) return fail("gamma must be >= 1 (got %d)", c->gamma);
    if (c->nu1 < 0 || c->nu2 < 0) return fail("nu1/nu2 must be >= 0");
    if (c->max_iters < 0) return fail("max_iters must be >= 0");
    return 0;
}

template <typename T>
Hier<T>& hier(jsv_plan* P);
template <>
Hier<double>& hier<double>(jsv_plan* P) { return P->h64; }
template <>
Hier<float>& hier<float>(jsv_plan* P) { return P->h32; }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int jsv_abi_version(void) { return 1; }

const char* jsv_last_error(void) { return g_err.c_str(); }

int jsv_plan_create_batched(int dim, int mx, int my, int mz, int batch, double idel_x,
                            double idel_y, double idel_z, int precision, jsv_plan** out)
{
    if (out == nullptr) return fail("out is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail("dim must be 2 or 3 (got %d)", dim);
    if (batch < 1 || (dim == 3 && batch != 1))
        return fail("batch must be >= 1 (2D) or 1 (3D), got %d", batch);
    if (precision != JSV_FP64 && precision != JSV_MIXED) return fail("unknown precision %d", precision);
    const int ms[3] = {mx, my, dim == 3 ? mz : 1};
    const char names[3] = {'x', 'y', 'z'};
    for (int a = 0; a < dim; ++a) {
        if (ms[a] < 2 || (ms[a] & (ms[a] - 1)) != 0)
            return fail("Resolution of %c-grid is not a power of 2. Value: %d", names[a], ms[a]);
    }
    int lv[3], L = 1 << 30;
    for (int a = 0; a < dim; ++a) {
        lv[a] = (int)std::lround(std::log2((double)ms[a])) + 1;
        L = lv[a] < L ? lv[a] : L;
    }
    for (int a = 0; a < dim; ++a) {
        const int coarsest = 1 << (lv[a] - L);
        if (coarsest > 5)
            return fail("Resolution of %c-grid not efficient for multigrid configuration. "
                        "Resolution of coarsest level: %d", names[a], coarsest);
    }
    if (L < 2) return fail("grid too small for multigrid (levels=%d)", L);
    jsv_plan* P = new jsv_plan();
    P->dim = dim;
    P->batch = batch;
    P->prec = precision;
    P->levels = L;
    // measured at 256^3: the FP64 cycle is fastest with the program from 8^3,
    // the FP32 cycle from 16^3
    P->coarse_m[0] = dim == 3 ? 8 : 64;
    P->coarse_m[1] = dim == 3 ? 16 : 64;
    P->coarse_lds = 1;
    // 3D FP32 restrictions of the 64^3 and 32^3 levels are faster in two passes
    P->two_pass_min[1] = dim == 3 ? (1L << 13) : (1L << 20);
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            P->n_cu = ncu;
    }
    for (int a = 0; a < 3; ++a) P->m[a] = ms[a];
    P->idel[0] = idel_x;
    P->idel[1] = idel_y;
    P->idel[2] = idel_z;
    int rc = alloc_hier(P, P->h64, true);   // D_xy arrays are added by set_operator
    if (rc == 0 && precision == JSV_MIXED) rc = alloc_hier(P, P->h32, true);
    const Geo& g0 = P->h64.g[0];
    if (rc == 0) rc = dalloc(P, (void**)&P->phi_old, sizeof(double) * g0.n);
    if (rc == 0) {
        P->npartials = (long)((g0.mx + jsv::BX - 1) / jsv::BX + 1) * ((g0.my + jsv::BY) / jsv::BY + 1) *
                       (dim == 3 ? g0.mz + 1 : batch);
        rc = dalloc(P, (void**)&P->partials, sizeof(double) * P->npartials);
    }
    if (rc == 0) rc = dalloc(P, (void**)&P->dscal, sizeof(double) * 4 * batch);
    if (rc == 0) rc = dalloc(P, (void**)&P->d_act, sizeof(int) * batch);
    if (rc == 0) rc = dalloc(P, (void**)&P->d_actk, sizeof(int) * ACTK_ROWS * batch);
    if (rc == 0) {
        hipError_t e = hipHostMalloc((void**)&P->hscal, sizeof(double) * 4 * batch, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&P->h_act, sizeof(int) * batch, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&P->h_actk, sizeof(int) * ACTK_ROWS * batch, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&P->h_stop, sizeof(jsv::StopCtl), hipHostMallocDefault);
        if (e != hipSuccess) rc = fail("hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    if (rc == 0) rc = dalloc(P, (void**)&P->d_stop, sizeof(jsv::StopCtl));
    if (rc != 0) {
        std::string keep = g_err;
        jsv_plan_destroy(P);
        g_err = keep;
        return rc;
    }
    *out = P;
    return 0;
}

int jsv_plan_create(int dim, int mx, int my, int mz, double idel_x, double idel_y, double idel_z,
                    int precision, jsv_plan** out)
{
    return jsv_plan_create_batched(dim, mx, my, mz, 1, idel_x, idel_y, idel_z, precision, out);
}

int jsv_plan_destroy(jsv_plan* P)
{
    if (P == nullptr) return 0;
    adj_forget(P);   // the adjoint state holds this plan's buffers
    for (auto& gr : P->graphs)
        if (gr.exec != nullptr) (void)hipGraphExecDestroy(gr.exec);   // nullptr: replayed eagerly
    if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
    for (void* p : P->allocs) (void)hipFree(p);
    if (P->hscal) (void)hipHostFree(P->hscal);
    if (P->h_act) (void)hipHostFree(P->h_act);
    if (P->h_actk) (void)hipHostFree(P->h_actk);
    if (P->h_stop) (void)hipHostFree(P->h_stop);
    delete P;
    return 0;
}

int jsv_plan_levels(const jsv_plan* P, int* levels)
{
    if (P == nullptr || levels == nullptr) return fail("NULL argument");
    *levels = P->levels;
    return 0;
}

int jsv_plan_level_shape(const jsv_plan* P, int level, int shape[3])
{
    if (P == nullptr || shape == nullptr) return fail("NULL argument");
    if (level < 0 || level >= P->levels) return fail("level %d out of range", level);
    const Geo& g = P->h64.g[level];
    shape[0] = g.nz;
    shape[1] = g.ny;
    shape[2] = g.nx;
    return 0;
}

int jsv_plan_set_tuning(jsv_plan* P, int key, int64_t value)
{
    if (P == nullptr) return fail("plan is NULL");
    if (value < 0) return fail("tuning values must be >= 0");
    drop_graphs(P);    // captured launch shapes depend on every knob
    if (key == JSV_TUNE_GRAPHS) {
        P->use_graphs = value != 0;
        return 0;
    }
    if (key == JSV_TUNE_SLAB_BYTES) {
        P->slab_bytes = (long)value;
        return 0;
    }
    if (key == JSV_TUNE_TWO_PASS_MIN) {
        P->two_pass_min[0] = P->two_pass_min[1] = (long)value;
        auto realloc_res = [&](auto& H) -> int {
            for (int l = 0; l + 1 < P->levels && l < (int)H.res.size(); ++l)
                if (H.res[l] == nullptr && H.g[l].n >= value)
                    RETCHK(dalloc(P, (void**)&H.res[l], sizeof(*H.phi[0]) * H.g[l].n));
            return 0;
        };
        RETCHK(realloc_res(P->h64));
        if (P->prec == JSV_MIXED) RETCHK(realloc_res(P->h32));
        // levels below the threshold go back to one pass
        for (int l = 0; l + 1 < P->levels; ++l) {
            if (P->h64.g[l].n < value) {
                if (l < (int)P->h64.res.size()) P->h64.res[l] = nullptr;
                if (l < (int)P->h32.res.size()) P->h32.res[l] = nullptr;
            }
        }
        return 0;
    }
    if (key == JSV_TUNE_FUSED_MIN_M) {
        P->fused_min_m = (int)value;
        auto realloc_phi2 = [&](auto& H) -> int {
            if (P->dim != 3) return 0;
            for (int l = 0; l < P->levels && l < (int)H.phi2.size(); ++l)
                if (H.phi2[l] == nullptr && value > 0 && H.g[l].mz >= value)
                    RETCHK(dalloc(P, (void**)&H.phi2[l], sizeof(*H.phi[0]) * H.g[l].n));
            return 0;
        };
        RETCHK(realloc_phi2(P->h64));
        if (P->prec == JSV_MIXED) RETCHK(realloc_phi2(P->h32));
        return 0;
    }
    if (key == JSV_TUNE_COARSE_M) {
        P->coarse_m[0] = P->coarse_m[1] = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_COARSE_LDS) {
        P->coarse_lds = value != 0 ? 1 : 0;
        P->lv_host[0].clear();      // restage on the next cycle
        P->lv_host[1].clear();
        return 0;
    }
    if (key == JSV_TUNE_COARSE_SOLO) {
        if (value > jsv::COARSE_WAVE) return fail("coarse solo limit must be <= %d points", jsv::COARSE_WAVE);
        P->coarse_solo = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_COARSE_THREADS) {
        if (value != 0 && value != 256 && value != 512 && value != 1024)
            return fail("coarse program threads must be 0, 256, 512 or 1024");
        P->coarse_threads = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_ZCHUNK) {
        P->zchunk = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_FUSE_HALVES) {
        P->fuse_halves = value != 0 ? 1 : 0;
        return 0;
    }
    if (key == JSV_TUNE_DEVICE_STOP) {
        P->device_stop = value != 0 ? 1 : 0;
        return 0;
    }
    if (key == JSV_TUNE_ZIGZAG) {
        if (value > 2) return fail("JSV_TUNE_ZIGZAG takes 0, 1 or 2, not %ld", (long)value);
        P->zigzag = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_FUSE_RESID) {
        P->fuse_resid = value != 0 ? 1 : 0;
        return 0;
    }
    if (key == JSV_TUNE_COARSE_REG) {
        if (value > 2) return fail("coarse reg must be 0, 1 or 2");
        P->coarse_reg = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_FUSE_PROLONG) {
        if (value > 2) return fail("fuse prolong must be 0, 1 or 2");
        P->fuse_prolong = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_SWEEP_PAIRS) {
        if (value > 2) return fail("sweep pairs must be 0, 1 or 2");
        P->sweep_pairs = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_SWEEP_VARIANT) {
        if (value > 2) return fail("sweep variant must be 0, 1 or 2");
        P->sweep_variant = value == 0 ? 2 : (int)value;
        return 0;
    }
    if (key == JSV_TUNE_SWEEP_ROWS) {
        if (value != 0 && value != 2 && value != 4 && value != 5 && value != 6 && value != 8 && value != 14 &&
            value != 16)
            return fail("sweep rows must be 0, 2, 4, 5, 6, 8, 14 or 16");
        P->sweep_rows = (int)value;
        return 0;
    }
    if (key == JSV_TUNE_SWEEP_THREADS) {
        if (value != 0 && value != 256 && value != 768 && value != 1024)
            return fail("sweep threads must be 0, 256, 768 or 1024");
        P->sweep_threads = (int)value;
        return 0;
    }
    return fail("unknown tuning key %d", key);
}

int jsv_plan_device_bytes(const jsv_plan* P, int64_t* bytes)
{
    if (P == nullptr || bytes == nullptr) return fail("NULL argument");
    *bytes = P->bytes;
    return 0;
}

int jsv_plan_set_operator(jsv_plan* P, const double* lam, double lam_scalar, const double* dxx,
                          const double* dyy, const double* dzz, const double d_scalar[3],
                          const double* dxy, void* stream)
{
    if (P == nullptr) return fail("plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int dim = P->dim;
    const bool spatial = dxx != nullptr;
    if (spatial && (dyy == nullptr || (dim == 3 && dzz == nullptr)))
        return fail("D components must all be arrays or all be scalars");
    if (!spatial && (dyy != nullptr || (dim == 3 && dzz != nullptr)))
        return fail("D components must all be arrays or all be scalars");
    if (!spatial && d_scalar == nullptr) return fail("scalar D values missing");
    if (dxy != nullptr && (dim != 2 || !spatial)) return fail("D_xy needs 2D array diffusion");
    P->spatial = spatial;
    P->lam_array = lam != nullptr;
    ++P->op_version;                                      // the coarse program's image is stale
    if (P->offdiag != (dxy != nullptr)) drop_graphs(P);   // the smoother's kernels change
    P->offdiag = dxy != nullptr;
    Hier<double>& H = P->h64;
    const int L = P->levels;
    if (P->offdiag && H.drr[0] == nullptr) {
        for (int l = 0; l < L; ++l) {
            RETCHK(dalloc(P, (void**)&H.drr[l], sizeof(double) * H.g[l].n));
            RETCHK(dalloc(P, (void**)&H.dlr[l], sizeof(double) * H.g[l].n));
        }
        if (P->prec == JSV_MIXED) {
            for (int l = 0; l < L; ++l) {
                RETCHK(dalloc(P, (void**)&P->h32.drr[l], sizeof(float) * H.g[l].n));
                RETCHK(dalloc(P, (void**)&P->h32.dlr[l], sizeof(float) * H.g[l].n));
            }
        }
    }
    const Geo& g0 = H.g[0];
    const double ds[3] = {d_scalar ? d_scalar[0] : 0.0, d_scalar ? d_scalar[1] : 0.0,
                          d_scalar ? d_scalar[2] : 0.0};
    if (dim == 2) {
        for (int l = 0; l < L; ++l) {
            const Geo& g = H.g[l];
            const double idx = P->idel[0] * std::ldexp(1.0, -l);
            const double idy = P->idel[1] * std::ldexp(1.0, -l);
            const double sqx = idx * idx, sqy = idy * idy;
            const double cxy = idx * idy * 0.25;
            const int stride = 1 << l;
            jsv::Src sDx{dxx, g0.sy, g0.sz, stride, l == 0};
            jsv::Src sDy{dyy, g0.sy, g0.sz, stride, l == 0};
            jsv::Src sLam{lam, g0.sy, g0.sz, stride, l == 0};
            jsv::Src sDxy{dxy, g0.sy, g0.sz, stride, l == 0};
            hipLaunchKernelGGL(jsv::k_setup2d, grid_exact(g.n), dim3(256), 0, s, H.cx[l], H.cy[l],
                               H.cc[l], P->offdiag ? H.drr[l] : nullptr,
                               P->offdiag ? H.dlr[l] : nullptr, g, sDx, sDy, sLam, sDxy, sqx, sqy,
                               cxy, lam_scalar, ds[0], ds[1], spatial ? 1 : 0);
            LAUNCHCHK();
        }
    } else {
        // averaged face arrays per level (level 0: the raw input)
        std::vector<double*> avg[3];
        std::vector<void*> tmp;
        int rc = 0;
        if (spatial) {
            for (int a = 0; a < 3 && rc == 0; ++a) {
                avg[a].assign(L, nullptr);
                avg[a][0] = const_cast<double*>(a == 0 ? dxx : a == 1 ? dyy : dzz);
                for (int l = 1; l < L && rc == 0; ++l) {
                    void* p = nullptr;
                    hipError_t e = hipMallocAsync(&p, sizeof(double) * H.g[l].n, s);
                    if (e != hipSuccess) rc = fail("hipMallocAsync failed: %s", hipGetErrorString(e));
                    avg[a][l] = (double*)p;
                    tmp.push_back(p);
                    if (rc == 0)
                        hipLaunchKernelGGL(jsv::k_avg3, grid_exact(H.g[l].n), dim3(256), 0, s,
                                           avg[a][l], avg[a][l - 1], H.g[l], H.g[l - 1], a);
                }
            }
        }
        for (int l = 0; l < L && rc == 0; ++l) {
            const Geo& g = H.g[l];
            const double idx = P->idel[0] * std::ldexp(1.0, -l);
            const double idy = P->idel[1] * std::ldexp(1.0, -l);
            const double idz = P->idel[2] * std::ldexp(1.0, -l);
            jsv::Src sLam{lam, g0.sy, g0.sz, 1 << l, l == 0};
            hipLaunchKernelGGL(jsv::k_setup3d, grid_exact(g.n), dim3(256), 0, s, H.cx[l], H.cy[l],
                               H.cz[l], H.cc[l], g, spatial ? avg[0][l] : nullptr,
                               spatial ? avg[1][l] : nullptr, spatial ? avg[2][l] : nullptr, sLam,
                               idx * idx, idy * idy, idz * idz, lam_scalar, ds[0], ds[1], ds[2],
                               spatial ? 1 : 0);
        }
        for (void* p : tmp) (void)hipFreeAsync(p, s);
        if (rc != 0) return rc;
        LAUNCHCHK();
    }
    if (P->prec == JSV_MIXED) {
        Hier<float>& F = P->h32;
        for (int l = 0; l < L; ++l) {
            const long n = H.g[l].n;
            auto cast = [&](const double* a, float* b) {
                if (a != nullptr && b != nullptr)
                    hipLaunchKernelGGL((jsv::k_cast<double, float>), grid_flat(n), dim3(256), 0, s,
                                       a, b, n);
            };
            cast(H.cx[l], F.cx[l]);
            cast(H.cy[l], F.cy[l]);
            if (dim == 3) cast(H.cz[l], F.cz[l]);
            cast(H.cc[l], F.cc[l]);
            if (P->offdiag) {
                cast(H.drr[l], F.drr[l]);
                cast(H.dlr[l], F.dlr[l]);
            }
        }
        LAUNCHCHK();
    }
    P->op_set = true;
    return 0;
}

int jsv_level_upload(jsv_plan* P, int prec, int level, int which, const void* src, void* stream)
{
    if (P == nullptr || src == nullptr) return fail("NULL argument");
    if (level < 0 || level >= P->levels) return fail("level %d out of range", level);
    if (prec == JSV_MIXED && P->prec != JSV_MIXED) return fail("plan has no FP32 hierarchy");
    const long n = P->h64.g[level].n;
    void* dst;
    size_t es;
    if (prec == JSV_MIXED) {
        dst = which == JSV_BUF_PHI ? (void*)P->h32.phi[level] : (void*)P->h32.rhs[level];
        es = sizeof(float);
    } else {
        dst = which == JSV_BUF_PHI ? (void*)P->h64.phi[level] : (void*)P->h64.rhs[level];
        es = sizeof(double);
    }
    HIPCHK(hipMemcpyAsync(dst, src, es * n, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return 0;
}

int jsv_level_download(jsv_plan* P, int prec, int level, int which, void* dst, void* stream)
{
    if (P == nullptr || dst == nullptr) return fail("NULL argument");
    if (level < 0 || level >= P->levels) return fail("level %d out of range", level);
    if (prec == JSV_MIXED && P->prec != JSV_MIXED) return fail("plan has no FP32 hierarchy");
    const long n = P->h64.g[level].n;
    const void* src;
    size_t es;
    if (prec == JSV_MIXED) {
        src = which == JSV_BUF_PHI ? (void*)P->h32.phi[level] : (void*)P->h32.rhs[level];
        es = sizeof(float);
    } else {
        src = which == JSV_BUF_PHI ? (void*)P->h64.phi[level] : (void*)P->h64.rhs[level];
        es = sizeof(double);
    }
    HIPCHK(hipMemcpyAsync(dst, src, es * n, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return 0;
}

int jsv_smooth(jsv_plan* P, int prec, int level, int sweeps, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (level < 0 || level >= P->levels) return fail("level %d out of range", level);
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < sweeps; ++i) {
        if (prec == JSV_MIXED) RETCHK(smooth(P, P->h32, level, s));
        else RETCHK(smooth(P, P->h64, level, s));
    }
    LAUNCHCHK();
    return 0;
}

int jsv_restrict(jsv_plan* P, int prec, int level, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (level < 0 || level >= P->levels - 1) return fail("level %d out of range", level);
    hipStream_t s = (hipStream_t)stream;
    if (prec == JSV_MIXED) RETCHK(restrict_(P, P->h32, level, s));
    else RETCHK(restrict_(P, P->h64, level, s));
    LAUNCHCHK();
    return 0;
}

int jsv_prolong_add(jsv_plan* P, int prec, int level, int nu2, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (level < 0 || level >= P->levels - 1) return fail("level %d out of range", level);
    hipStream_t s = (hipStream_t)stream;
    if (prec == JSV_MIXED) RETCHK(prolong(P, P->h32, level, nu2, s));
    else RETCHK(prolong(P, P->h64, level, nu2, s));
    LAUNCHCHK();
    return 0;
}

int jsv_cycle(jsv_plan* P, int prec, const jsv_config* cfg, int skip_pre, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    RETCHK(check_cfg(cfg));
    hipStream_t s = (hipStream_t)stream;
    if (prec == JSV_MIXED) {
        if (P->prec != JSV_MIXED) return fail("plan has no FP32 hierarchy");
        // rim of the FP32 level-0 phi: zero via the same kernel on a temp view is
        // not available for float; the FP32 buffers are only written on unknowns
        RETCHK(prepare_coarse(P, P->h32, *cfg, s));
        RETCHK(mg_cycle(P, P->h32, *cfg, 0, cfg->fcycle != 0, skip_pre != 0, s));
    } else {
        RETCHK(zero_rim_level0(P, P->h64.phi[0], s));
        RETCHK(prepare_coarse(P, P->h64, *cfg, s));
        RETCHK(mg_cycle(P, P->h64, *cfg, 0, cfg->fcycle != 0, skip_pre != 0, s));
    }
    LAUNCHCHK();
    return 0;
}

int jsv_residual(jsv_plan* P, const double* phi, const double* rhs, double* out_r,
                 double* norm2_out, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (phi == nullptr || rhs == nullptr) return fail("NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const long n = P->h64.g[0].n;
    if (out_r != nullptr) RETCHK(zero(out_r, n, s));
    // phi's boundary nodes and ghosts are the zero Dirichlet values: use a copy
    // with the rim cleared, as the solver's own iterates always have it
    HIPCHK(hipMemcpyAsync(P->phi_old, phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RETCHK(zero_rim_level0(P, P->phi_old, s));
    RETCHK(residual64<double>(P, P->phi_old, rhs, out_r, 0, s));
    LAUNCHCHK();
    if (norm2_out != nullptr) {
        RETCHK(read_scalars(P, 1, s));
        for (int b = 0; b < nseg(P); ++b) norm2_out[b] = P->hscal[b];
    }
    return 0;
}

int jsv_dirichlet_lift(jsv_plan* P, const double* phi_b, const double* rhs, double* rhs_out, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (phi_b == nullptr || rhs == nullptr || rhs_out == nullptr) return fail("NULL argument");
    const Hier<double>& H = P->h64;
    const Geo& g = H.g[0];
    const bool od = P->dim == 2 && P->offdiag;
    hipLaunchKernelGGL(jsv::k_dirichlet_lift, grid_exact(g.n), dim3(256), 0, (hipStream_t)stream, rhs_out, rhs,
                       phi_b, H.cx[0], H.cy[0], P->dim == 3 ? H.cz[0] : nullptr, od ? H.drr[0] : nullptr,
                       od ? H.dlr[0] : nullptr, g, P->dim);
    LAUNCHCHK();
    return 0;
}

int jsv_max_abs_diff(const double* a, const double* b, int64_t n, double* out, void* stream)
{
    if (a == nullptr || b == nullptr || out == nullptr) return fail("NULL argument");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* d = nullptr;
    HIPCHK(hipMallocAsync((void**)&d, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(jsv::k_set_u64, dim3(1), dim3(64), 0, s, d, 0ull);
    long nb = (long)((n + 255) / 256);
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(jsv::k_maxabsdiff<double>, dim3((unsigned)nb), dim3(256), 0, s, a, b,
                       (long)n, d);
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFreeAsync(d, s);
    if (e != hipSuccess) return fail("max_abs_diff failed: %s", hipGetErrorString(e));
    double v;
    std::memcpy(&v, &h, sizeof(v));
    *out = v;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// jsv_max_abs_diff without allocation: its result words live in a
// file-scope device array (each GPU holds its own copy until the process
// ends); the host reaches it through the address hipGetSymbolAddress returns
// on that device, looked up once per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[JSV_NOALLOC_MAX_SEGMENTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// this device's address of `scratch` (the current device)
int scratch_ptr(unsigned long long** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail("device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail("hipGetSymbolAddress returned no address");
        g_scratch_ptr[dev] = (unsigned long long*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int jsv_max_abs_diff_noalloc(const double* a, const double* b, int64_t n, int64_t nseg, double* out,
                             void* stream)
{
    if (a == nullptr || b == nullptr || out == nullptr) return fail("NULL argument");
    if (n < 0) return fail("n must be >= 0 (got %lld)", (long long)n);
    if (nseg < 1 || nseg > JSV_NOALLOC_MAX_SEGMENTS)
        return fail("jsv_max_abs_diff_noalloc: %lld segments requested, the scratch holds at most %d",
                    (long long)nseg, JSV_NOALLOC_MAX_SEGMENTS);
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* d = nullptr;
    RETCHK(scratch_ptr(&d));
    // the words keep the last call's maxima: start from 0 (a fill kernel, no memset)
    hipLaunchKernelGGL(jsv::k_fill<unsigned long long>, dim3(1), dim3(256), 0, s, d, (long)nseg, 0ull);
    long nb = (long)((n + 255) / 256);
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(jsv::k_maxabsdiff<double>, dim3((unsigned)nb, (unsigned)nseg), dim3(256), 0, s, a, b,
                       (long)n, d);
    LAUNCHCHK();
    // non-negative doubles travel as their bit patterns
    static_assert(sizeof(double) == sizeof(unsigned long long), "64-bit words");
    HIPCHK(hipMemcpyAsync(out, d, sizeof(double) * (size_t)nseg, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int jsv_stream_bench(const float* src, int k, int64_t n, float* dst, int blocks, void* stream)
{
    if (src == nullptr || dst == nullptr) return fail("NULL argument");
    if (n <= 0 || n % 4 != 0) return fail("n must be a positive multiple of 4 (got %lld)", (long long)n);
    if (((uintptr_t)src | (uintptr_t)dst) % 16 != 0) return fail("arrays must be 16-byte aligned");
    const long n4 = (long)(n / 4);
    long nb = blocks > 0 ? blocks : (n4 + 255) / 256;
    if (nb > 65535L * 16) nb = 65535L * 16;
    const jsv::f32x4* s4 = reinterpret_cast<const jsv::f32x4*>(src);
    jsv::f32x4* d4 = reinterpret_cast<jsv::f32x4*>(dst);
    hipStream_t s = (hipStream_t)stream;
    switch (k) {
    case 1: hipLaunchKernelGGL(jsv::k_stream<1>, dim3((unsigned)nb), dim3(256), 0, s, s4, d4, n4); break;
    case 2: hipLaunchKernelGGL(jsv::k_stream<2>, dim3((unsigned)nb), dim3(256), 0, s, s4, d4, n4); break;
    case 6: hipLaunchKernelGGL(jsv::k_stream<6>, dim3((unsigned)nb), dim3(256), 0, s, s4, d4, n4); break;
    default: return fail("k must be 1, 2 or 6 (got %d)", k);
    }
    LAUNCHCHK();
    return 0;
}

static int do_bcs(double* arr, int dim, int nz, int ny, int nx, int rim, int direction,
                  const double* vxb, const double* vxe, const double* vyb, const double* vye,
                  void* stream)
{
    if (arr == nullptr) return fail("NULL array");
    if (rim < 0) return fail("rim must be >= 0");
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.sy = nx;
    g.sz = (long)nx * ny;
    g.n = g.sz * nz;
    hipStream_t s = (hipStream_t)stream;
    for (int ph = 0; ph < dim; ++ph)
        hipLaunchKernelGGL(jsv::k_bcs, grid_exact(g.n), dim3(256), 0, s, arr, g, dim, rim,
                           direction, vxb, vxe, vyb, vye, ph);
    LAUNCHCHK();
    return 0;
}

int jsv_do_bcs_2d(double* arr, int ny, int nx, int rim, int direction, const double* vals_xb,
                  const double* vals_xe, const double* vals_yb, const double* vals_ye, void* stream)
{
    return do_bcs(arr, 2, 1, ny, nx, rim, direction, vals_xb, vals_xe, vals_yb, vals_ye, stream);
}

int jsv_do_bcs_3d(double* arr, int nz, int ny, int nx, int rim, int direction,
                  const double* vals_xb, const double* vals_xe, const double* vals_yb,
                  const double* vals_ye, void* stream)
{
    return do_bcs(arr, 3, nz, ny, nx, rim, direction, vals_xb, vals_xe, vals_yb, vals_ye, stream);
}

// ---------------------------------------------------------------------------
// solve loop (Solver_2D.solve, solver_2d.py:423-479)
// ---------------------------------------------------------------------------
static void record(const jsv_config* c, int64_t k, double v)
{
    if (c->history != nullptr && k < c->history_len) c->history[k] = v;
}

// Per-problem bookkeeping of a (batched) dynamic solve.  A problem stops
// after the first cycle whose stop quantity is not above the threshold (NaN
// stops it too, as `dist > eps` is false for NaN in the reference); kernels of
// later cycles skip it through the device flags.
// Cycles until the relative residual is predicted to reach rtol, from the
// contraction measured over the last k_last cycles (rel_prev -> rel_now);
// clamped to [1, kmax].  Used to space residual checks (FP64, check_every = 0)
// and FP64 refinement steps (mixed, inner_cycles = 0).
static int64_t predict_cycles(double rel_prev, double rel_now, int64_t k_last, double rtol,
                              int64_t kfirst, int64_t kmax)
{
    if (k_last <= 0 || !(rel_prev > 0.0) || !(rel_now > 0.0) || !std::isfinite(rel_now)) return kfirst;
    const double rate = std::pow(rel_now / rel_prev, 1.0 / (double)k_last);
    if (!(rate < 0.999999)) return kmax;
    const double need = std::log(rtol / rel_now) / std::log(rate);
    if (!(need > 1.0)) return 1;
    const double k = std::ceil(need);
    return k > (double)kmax ? kmax : (int64_t)k;
}

struct Track {
    std::vector<int64_t> count, outer;
    std::vector<double> last, rel;
    std::vector<int> nonfinite;
    int active = 0;
    void init(int B)
    {
        count.assign(B, 0);
        outer.assign(B, 0);
        last.assign(B, INFINITY);
        rel.assign(B, -1.0);
        nonfinite.assign(B, 0);
        active = B;
    }
};

static int push_act(jsv_plan* P, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(P->d_act, P->h_act, sizeof(int) * P->batch, hipMemcpyHostToDevice, s));
    return 0;
}

// Runs `rows` cycles in which problem b takes part in the first kb[b] only
// (its own schedule, so a batched solve gives each problem the bits of its
// single solve).  Batched plans stage one activity row per cycle and copy a
// row into the flags the kernels read whenever it changes; the outer flags
// are restored afterwards.
extern "C++" template <typename T>
static int run_cycles_sched(jsv_plan* P, Hier<T>& H, const jsv_config& c, bool skip_pre,
                           const std::vector<int64_t>& kb, int64_t rows, hipStream_t s)
{
    const int B = P->batch;
    if (!P->use_act) {
        for (int64_t j = 0; j < rows; ++j) RETCHK(run_cycle(P, H, c, skip_pre, s));
        return 0;
    }
    for (int64_t j0 = 0; j0 < rows; j0 += ACTK_ROWS) {
        const int64_t nr = rows - j0 < ACTK_ROWS ? rows - j0 : ACTK_ROWS;
        if (j0 > 0) HIPCHK(hipStreamSynchronize(s));   // the staging buffer is reused
        for (int64_t j = 0; j < nr; ++j)
            for (int b = 0; b < B; ++b)
                P->h_actk[j * B + b] = (P->h_act[b] && j0 + j < kb[b]) ? 1 : 0;
        HIPCHK(hipMemcpyAsync(P->d_actk, P->h_actk, sizeof(int) * nr * B, hipMemcpyHostToDevice, s));
        for (int64_t j = 0; j < nr; ++j) {
            const bool same = (j > 0) && std::memcmp(P->h_actk + j * B, P->h_actk + (j - 1) * B,
                                                     sizeof(int) * B) == 0;
            if (!same)
                HIPCHK(hipMemcpyAsync(P->d_act, P->d_actk + j * B, sizeof(int) * B,
                                      hipMemcpyDeviceToDevice, s));
            RETCHK(run_cycle(P, H, c, skip_pre, s));
        }
    }
    return push_act(P, s);
}

static void fill_result(const Track& T, jsv_result* res, int64_t* counts, double* stopvals, int B)
{
    int64_t mx = 0, mo = 0;
    int nf = 0;
    for (int b = 0; b < B; ++b) {
        mx = T.count[b] > mx ? T.count[b] : mx;
        mo = T.outer[b] > mo ? T.outer[b] : mo;
        nf |= T.nonfinite[b];
        if (counts) counts[b] = T.count[b];
        if (stopvals) stopvals[b] = T.last[b];
    }
    res->count = B == 1 ? T.count[0] : mx;
    res->outer_steps = B == 1 ? T.outer[0] : mo;
    res->nonfinite = nf;
    res->rel_residual = T.rel[0];
}

// A dynamic solve of one problem that tests every cycle (the reference's
// distance rule, or the relative residual with check_every = 1) runs its test
// on the device (k_stop_check inside each cycle): the host queues cycles in
// blocks and reads the control block once per block; cycles queued behind a
// stop do nothing.  The block length comes from the stop values' measured
// contraction, so the last block ends near the stop.  Count, stop values and
// phi are those of the test after every cycle.
static bool device_stop_applies(const jsv_plan* P, const jsv_config* c, bool relres, int B)
{
    return P->device_stop != 0 && B == 1 && !c->fixed && c->max_iters > 0 && (!relres || c->check_every == 1);
}

static int solve_device_stop(jsv_plan* P, const jsv_config& c, bool relres, double rhs_norm,
                             const double* first_old, Track& T, hipStream_t s)
{
    Hier<double>& H = P->h64;
    const long n = H.g[0].n;
    jsv::StopCtl& h = *P->h_stop;
    long hl = 0;      // stop values kept on the device for the history
    if (c.history != nullptr && c.history_len > 0) {
        hl = (long)std::min<int64_t>(c.history_len, c.max_iters);
        if (hl > P->hist_cap) {
            HIPCHK(hipStreamSynchronize(s));
            RETCHK(dfree(P, P->d_hist, sizeof(double) * (size_t)P->hist_cap));
            P->d_hist = nullptr;
            P->hist_cap = 0;
            RETCHK(dalloc(P, (void**)&P->d_hist, sizeof(double) * (size_t)hl));
            P->hist_cap = hl;
        }
    }
    h = jsv::StopCtl{relres ? c.rtol : c.epsilon, rhs_norm, INFINITY, (long long)c.max_iters, 0, 0, 0, 0,
                     hl > 0 ? P->d_hist : nullptr, (long long)hl};
    HIPCHK(hipMemcpyAsync(P->d_stop, P->h_stop, sizeof(jsv::StopCtl), hipMemcpyHostToDevice, s));
    // the distance of the first cycle is from the caller's phi
    if (!relres && first_old != P->phi_old)
        HIPCHK(hipMemcpyAsync(P->phi_old, first_old, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    P->use_act = true;
    P->stop_mode = relres ? 1 : 2;
    // a relative-residual test runs at the start of the next cycle
    const int64_t cap = relres ? c.max_iters + 1 : c.max_iters;
    std::vector<double*> starts;     // the level-0 buffer each queued cycle starts from
    int64_t launched = 0, k = std::min<int64_t>(8, cap);
    double v_prev = -1.0;
    int64_t n_prev = 0;
    int rc = 0;
    for (;;) {
        for (int64_t j = 0; j < k && rc == 0; ++j) {
            starts.push_back(H.phi[0]);
            rc = run_cycle(P, H, c, c.skip_pre != 0, s);
        }
        if (rc != 0) break;
        launched += k;
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(P->h_stop, P->d_stop, sizeof(jsv::StopCtl), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h.stopped) break;
        if (launched >= cap) {
            rc = fail("internal: the device stop test did not end the solve");
            break;
        }
        int64_t kb = 8;
        if (h.count > 0) kb = predict_cycles(v_prev, h.last, h.count - n_prev, h.thr, 8, 64);
        if (relres) ++kb;            // the cycle whose start tests the last one
        v_prev = h.last;
        n_prev = h.count;
        k = std::min<int64_t>(kb, cap - launched);
    }
    P->stop_mode = 0;
    P->use_act = false;
    if (rc != 0) return rc;
    const int64_t cnt = h.count;
    // phi after cnt cycles: the buffer the next queued cycle started from
    // (cycles run eagerly may leave level 0's buffers swapped)
    double* out = cnt < (int64_t)starts.size() ? starts[cnt] : H.phi[0];
    if (out != H.phi[0]) {
        if (out != H.phi2[0]) return fail("internal: result of the device stop test not found");
        std::swap(H.phi[0], H.phi2[0]);
    }
    T.count[0] = T.outer[0] = cnt;
    T.last[0] = h.last;
    T.rel[0] = relres ? h.last : -1.0;
    T.nonfinite[0] = h.nonfinite;
    T.active = 0;
    const long nh = std::min<long>((long)cnt, hl);
    if (nh > 0) HIPCHK(hipMemcpy(c.history, P->d_hist, sizeof(double) * (size_t)nh, hipMemcpyDeviceToHost));
    return 0;
}

static int solve_fp64(jsv_plan* P, const double* phi_in, const double* rhs, double* phi_out,
                      const jsv_config* c, jsv_result* res, int64_t* counts, double* stopvals,
                      hipStream_t s)
{
    Hier<double>& H = P->h64;
    const long n = H.g[0].n;
    const int B = nseg(P);
    HIPCHK(hipMemcpyAsync(H.rhs[0], rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(H.phi[0], phi_in, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RETCHK(zero_rim_level0(P, H.phi[0], s));
    const bool relres = !c->fixed && c->stop == JSV_STOP_REL_RESIDUAL;
    std::vector<double> rhs_norm(B, 0.0);
    if (relres) {
        RETCHK(norm2_unknowns<double>(P, H.rhs[0], 1, s));
        RETCHK(read_scalars(P, 2, s));
        for (int b = 0; b < B; ++b) rhs_norm[b] = std::sqrt(P->hscal[B + b]);
    }
    Track T;
    T.init(B);
    for (int b = 0; b < B; ++b) P->h_act[b] = 1;
    RETCHK(push_act(P, s));
    P->use_act = !c->fixed && B > 1;
    // reference: initial skip_pre smoothing on the incoming phi
    const double* first_old = phi_in;
    if (c->skip_pre) {
        for (int i = 0; i < c->nu1; ++i) RETCHK(smooth(P, H, 0, s));
        HIPCHK(hipMemcpyAsync(P->phi_old, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        first_old = P->phi_old;
    }
    int rc = 0;
    if (c->fixed) {
        for (int64_t k = 0; k < c->max_iters && rc == 0; ++k) rc = run_cycle(P, H, *c, c->skip_pre != 0, s);
        for (int b = 0; b < B; ++b) T.count[b] = T.outer[b] = c->max_iters;
    } else if (device_stop_applies(P, c, relres, B)) {
        rc = solve_device_stop(P, *c, relres, relres ? rhs_norm[0] : 1.0, first_old, T, s);
    } else {
        const double thr = relres ? c->rtol : c->epsilon;
        for (int b = 0; b < B; ++b)
            if (c->max_iters <= 0) { P->h_act[b] = 0; --T.active; }
        if (T.active < B) RETCHK(push_act(P, s));
        // residual checks every cycle (reference-like exact count) or spaced by
        // the predicted cycles-to-tolerance (check_every = 0, rel_residual
        // only).  Each problem keeps its own check schedule: every active
        // problem runs every cycle, and at a check only the problems due are
        // tested against the tolerance, as in their single solves.
        const bool adaptive = relres && c->check_every == 0;
        const int64_t every = relres ? (c->check_every > 0 ? c->check_every : 1) : 1;
        const int64_t first_block = adaptive ? 8 : every;
        std::vector<int64_t> due(B), last_chk(B, 0), nchk(B, 0);
        std::vector<double> rel_prev(B, -1.0);
        for (int b = 0; b < B; ++b) due[b] = first_block < c->max_iters ? first_block : c->max_iters;
        int64_t done = 0;
        while (T.active > 0 && rc == 0) {
            const double* old = (done == 0) ? first_old : P->phi_old;
            if (!relres && done > 0)
                HIPCHK(hipMemcpyAsync(P->phi_old, H.phi[0], sizeof(double) * n,
                                      hipMemcpyDeviceToDevice, s));
            int64_t next = INT64_MAX;
            for (int b = 0; b < B; ++b)
                if (P->h_act[b] && due[b] < next) next = due[b];
            int64_t k = next - done;
            if (k < 1) k = 1;
            for (int64_t j = 0; j < k && rc == 0; ++j) rc = run_cycle(P, H, *c, c->skip_pre != 0, s);
            if (rc) break;
            done += k;
            LAUNCHCHK();
            if (relres) RETCHK(residual64<double>(P, H.phi[0], H.rhs[0], (double*)nullptr, 0, s));
            else RETCHK(maxdiff(P, old, H.phi[0], 0, s));
            RETCHK(read_scalars(P, 1, s));
            bool changed = false;
            for (int b = 0; b < B; ++b) {
                if (!P->h_act[b] || due[b] > done) continue;
                const double v = relres ? std::sqrt(P->hscal[b]) / rhs_norm[b] : P->hscal[b];
                T.last[b] = v;
                if (relres) T.rel[b] = v;
                if (!std::isfinite(v)) T.nonfinite[b] = 1;
                T.count[b] = T.outer[b] = done;
                if (b == 0) record(c, done - 1, v);
                if (!(v > thr) || done >= c->max_iters) {
                    P->h_act[b] = 0;
                    --T.active;
                    changed = true;
                    continue;
                }
                int64_t kb = every;
                if (adaptive) kb = predict_cycles(rel_prev[b], v, nchk[b] == 0 ? 0 : done - last_chk[b], thr, 8, 64);
                rel_prev[b] = v;
                last_chk[b] = done;
                ++nchk[b];
                due[b] = done + kb < c->max_iters ? done + kb : c->max_iters;
            }
            if (changed && T.active > 0) RETCHK(push_act(P, s));
        }
    }
    P->use_act = false;
    if (rc) return rc;
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(phi_out, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    fill_result(T, res, counts, stopvals, B);
    res->distance = relres ? -1.0 : T.last[0];
    return 0;
}

// FP32 cycles inside FP64 iterative refinement
static int solve_mixed(jsv_plan* P, const double* phi_in, const double* rhs, double* phi_out,
                       const jsv_config* c, jsv_result* res, int64_t* counts, double* stopvals,
                       hipStream_t s)
{
    Hier<double>& H = P->h64;
    Hier<float>& F = P->h32;
    const Geo& g = H.g[0];
    const long n = g.n;
    const int B = nseg(P);
    // FP32 cycles per refinement step: fixed, or (inner_cycles = 0) predicted
    // from the measured contraction so steps land near the tolerance
    // per problem, as in its single solve
    const bool adaptive = c->inner_cycles <= 0;
    std::vector<double> rel_prev(B, -1.0);
    std::vector<int64_t> k_last(B, 0), kb(B, 0);
    HIPCHK(hipMemcpyAsync(H.rhs[0], rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(H.phi[0], phi_in, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RETCHK(zero_rim_level0(P, H.phi[0], s));
    RETCHK(norm2_unknowns<double>(P, H.rhs[0], 1, s));
    RETCHK(read_scalars(P, 2, s));
    std::vector<double> rhs_norm(B);
    for (int b = 0; b < B; ++b) rhs_norm[b] = std::sqrt(P->hscal[B + b]);
    const bool relres = c->fixed || c->stop == JSV_STOP_REL_RESIDUAL;
    const double thr = c->fixed ? -1.0 : (relres ? c->rtol : c->epsilon);
    dim3 gr = grid_points(g, P->dim, g.mx - 1);
    Track T;
    T.init(B);
    for (int b = 0; b < B; ++b) P->h_act[b] = 1;
    RETCHK(push_act(P, s));
    P->use_act = B > 1;
    int rc = 0;
    bool first = true;
    while (rc == 0) {
        // FP64 residual, stored as the FP32 right-hand side of the correction equation
        RETCHK(residual64<float>(P, H.phi[0], H.rhs[0], F.rhs[0], 0, s));
        RETCHK(read_scalars(P, 1, s));
        bool changed = false;
        for (int b = 0; b < B; ++b) {
            if (!P->h_act[b]) continue;
            const double rel = std::sqrt(P->hscal[b]) / rhs_norm[b];
            T.rel[b] = rel;
            if (!std::isfinite(rel)) T.nonfinite[b] = 1;
            if (relres) {
                T.last[b] = rel;
                if (b == 0) record(c, T.outer[0], rel);
            }
            const bool stop = relres ? !(rel > thr) : (!first && !(T.last[b] > thr));
            if (stop || T.count[b] >= c->max_iters) {
                P->h_act[b] = 0;
                --T.active;
                changed = true;
            }
        }
        first = false;
        if (T.active == 0) break;
        if (changed) RETCHK(push_act(P, s));
        int64_t todo = 0;
        for (int b = 0; b < B; ++b) {
            kb[b] = 0;
            if (!P->h_act[b]) continue;
            int64_t k = c->inner_cycles;
            // fixed mode has no tolerance: a step aims at the five decades an
            // FP32 correction resolves
            const double tgt = c->fixed ? T.rel[b] * 1e-5 : thr;
            if (adaptive) k = relres ? predict_cycles(rel_prev[b], T.rel[b], k_last[b], tgt, 8, 64) : 8;
            if (T.count[b] + k > c->max_iters) k = c->max_iters - T.count[b];
            kb[b] = k;
            todo = k > todo ? k : todo;
        }
        for (int b = 0; b < B; ++b) rel_prev[b] = T.rel[b];
        // inner FP32 cycles on the correction, from a zero start
        RETCHK(zero(F.phi[0], n, s));
        rc = run_cycles_sched(P, F, *c, false, kb, todo, s);
        if (rc) break;
        if (!relres)
            HIPCHK(hipMemcpyAsync(P->phi_old, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(jsv::k_correct<float>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[0], F.phi[0],
                           g, P->dim, P->act_ptr());
        LAUNCHCHK();
        for (int b = 0; b < B; ++b)
            if (P->h_act[b]) {
                T.count[b] += kb[b];
                k_last[b] = kb[b];
                ++T.outer[b];
            }
        if (!relres) {
            RETCHK(maxdiff(P, P->phi_old, H.phi[0], 0, s));
            RETCHK(read_scalars(P, 1, s));
            for (int b = 0; b < B; ++b) {
                if (!P->h_act[b]) continue;
                T.last[b] = P->hscal[b];
                if (!std::isfinite(T.last[b])) T.nonfinite[b] = 1;
                if (b == 0) record(c, T.outer[0] - 1, T.last[0]);
            }
        }
    }
    P->use_act = false;
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(phi_out, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    fill_result(T, res, counts, stopvals, B);
    res->distance = relres ? -1.0 : T.last[0];
    return 0;
}

int jsv_solve_batched(jsv_plan* P, const double* phi_in, const double* rhs, double* phi_out,
                      const jsv_config* cfg, jsv_result* res, int64_t* counts, double* stopvals,
                      void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (phi_in == nullptr || rhs == nullptr || phi_out == nullptr || res == nullptr)
        return fail("NULL argument");
    RETCHK(check_cfg(cfg));
    hipStream_t s = (hipStream_t)stream;
    if (P->prec == JSV_MIXED) return solve_mixed(P, phi_in, rhs, phi_out, cfg, res, counts, stopvals, s);
    return solve_fp64(P, phi_in, rhs, phi_out, cfg, res, counts, stopvals, s);
}

int jsv_solve(jsv_plan* P, const double* phi_in, const double* rhs, double* phi_out,
              const jsv_config* cfg, jsv_result* res, void* stream)
{
    return jsv_solve_batched(P, phi_in, rhs, phi_out, cfg, res, nullptr, nullptr, stream);
}

}  // extern "C"

// ===========================================================================
// Reverse pass of fixed-mode solves (FP64)
// ===========================================================================
#include "jsv_adjoint.hpp"

namespace {

struct Adj {
    bool ready = false;
    int64_t bytes = 0;
    std::vector<double*> g, t, rb, rbar, bx, by, bz, bc, chain;
    std::vector<double*> brr, blr;            // 2D D_xy coupling bars (allocated on first use)
    double* sums = nullptr;                   // scalar cotangents, device double[4]
    std::vector<std::vector<double*>> pool;   // per level snapshot slots
    std::vector<size_t> used;
    std::vector<std::pair<int, double*>> tape;
    // one reversed cycle (recompute with tape + transposed cycle) per
    // configuration, captured as a hipGraph: every buffer it touches belongs
    // to the plan, and a cycle's input arrives in phi of level 0
    struct Graph {
        int gamma, fcycle, skip_pre, nu1, nu2;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};

// snapshots per level for one cycle (mirrors mg_cycle)
// snapshots per sweep: one per RBGS half, or one per sub-colour with D_xy
long ops_per_sweep(const jsv_plan* P) { return (P->dim == 2 && P->offdiag) ? 4L : 2L; }

void count_tape(const jsv_plan* P, const jsv_config& c, int l, bool fcyc, bool skip_pre,
                std::vector<long>& cnt)
{
    const long w = ops_per_sweep(P);
    if (l == P->levels - 2) {
        cnt[l] += w * (c.nu1 + c.nu2);
        return;
    }
    if (!skip_pre) cnt[l] += w * c.nu1;
    cnt[l] += 1;   // before the restriction
    if (fcyc)
        for (int i = 0; i < c.gamma; ++i) count_tape(P, c, l + 1, true, false, cnt);
    for (int i = 0; i < c.gamma; ++i) count_tape(P, c, l + 1, false, false, cnt);
    cnt[l] += w * c.nu2;
}

// one device array of every level's size
int one_level_set(jsv_plan* P, Adj& A, std::vector<double*>& v)
{
    const Hier<double>& H = P->h64;
    v.assign(P->levels, nullptr);
    for (int l = 0; l < P->levels; ++l) {
        RETCHK(dalloc(P, (void**)&v[l], sizeof(double) * H.g[l].n));
        A.bytes += (int64_t)sizeof(double) * H.g[l].n;
    }
    return 0;
}

int adj_alloc(jsv_plan* P, Adj& A, const jsv_config& c)
{
    const int L = P->levels;
    const Hier<double>& H = P->h64;
    if (!A.ready) {
        auto one = [&](std::vector<double*>& v, bool need) -> int {
            v.assign(L, nullptr);
            if (!need) return 0;
            for (int l = 0; l < L; ++l) {
                RETCHK(dalloc(P, (void**)&v[l], sizeof(double) * H.g[l].n));
                A.bytes += (int64_t)sizeof(double) * H.g[l].n;
            }
            return 0;
        };
        RETCHK(one(A.g, true));
        RETCHK(one(A.t, true));
        RETCHK(one(A.rb, true));
        RETCHK(one(A.rbar, true));
        RETCHK(one(A.bx, true));
        RETCHK(one(A.by, true));
        RETCHK(one(A.bz, P->dim == 3));
        RETCHK(one(A.bc, true));
        RETCHK(one(A.chain, true));
        A.pool.assign(L, {});
        RETCHK(dalloc(P, (void**)&A.sums, sizeof(double) * 4));
        A.bytes += (int64_t)sizeof(double) * 4;
        A.ready = true;
    }
    if (P->dim == 2 && P->offdiag && A.brr.empty()) {
        RETCHK(one_level_set(P, A, A.brr));
        RETCHK(one_level_set(P, A, A.blr));
    }
    std::vector<long> cnt(L, 0);
    count_tape(P, c, 0, c.fcycle != 0, c.skip_pre != 0, cnt);
    // the optional pre-loop smoothing (skip_pre) also records on level 0
    if (c.skip_pre) cnt[0] = cnt[0] > ops_per_sweep(P) * c.nu1 ? cnt[0] : ops_per_sweep(P) * c.nu1;
    for (int l = 0; l < L; ++l) {
        while ((long)A.pool[l].size() < cnt[l]) {
            double* p = nullptr;
            RETCHK(dalloc(P, (void**)&p, sizeof(double) * H.g[l].n));
            A.bytes += (int64_t)sizeof(double) * H.g[l].n;
            A.pool[l].push_back(p);
        }
    }
    A.used.assign(L, 0);
    A.tape.clear();
    return 0;
}

int snap(jsv_plan* P, Adj& A, int l, hipStream_t s)
{
    if (A.used[l] >= A.pool[l].size()) return fail("adjoint tape overflow at level %d", l);
    double* dst = A.pool[l][A.used[l]++];
    // a copy kernel, not a memcpy node, so the reversed cycle can be captured
    const long n = P->h64.g[l].n;
    hipLaunchKernelGGL((jsv::k_cast<double, double>), grid_flat(n), dim3(256), 0, s, P->h64.phi[l], dst, n);
    A.tape.push_back({l, dst});
    return 0;
}

int first_parity(const jsv_plan* P) { return P->dim == 3 ? 1 : 0; }

int rec_half(jsv_plan* P, Adj& A, int l, int par, hipStream_t s)
{
    // the half and its snapshot in one launch
    Hier<double>& H = P->h64;
    const Geo& g = H.g[l];
    if (A.used[l] >= A.pool[l].size()) return fail("adjoint tape overflow at level %d", l);
    double* dst = A.pool[l][A.used[l]++];
    const long n = P->dim == 3 ? g.n : g.sz;
    hipLaunchKernelGGL(jsv::k_rb_snap, grid_exact(n), dim3(256), 0, s, H.phi[l], dst, H.rhs[l], H.cx[l], H.cy[l],
                       P->dim == 3 ? H.cz[l] : H.cx[l], H.cc[l], g, par, P->dim);
    A.tape.push_back({l, dst});
    return 0;
}

// the reference's D_xy sub-colour order (solver_2d.py:686)
constexpr int kDiagOrder[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};

int rec_smooth(jsv_plan* P, Adj& A, int l, hipStream_t s)
{
    if (P->dim == 2 && P->offdiag) {
        // each sub-colour, then its snapshot
        Hier<double>& H = P->h64;
        const Geo& g = H.g[l];
        const dim3 gr((unsigned)((g.mx / 2 + jsv::BX - 1) / jsv::BX), (unsigned)((g.my / 2 + jsv::BY - 1) / jsv::BY), 1);
        for (const auto& o : kDiagOrder) {
            hipLaunchKernelGGL(jsv::k_rb2_diag<double>, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], H.rhs[l],
                               H.cx[l], H.cy[l], H.cc[l], H.drr[l], H.dlr[l], g, o[0], o[1], (const int*)nullptr);
            RETCHK(snap(P, A, l, s));
        }
        return 0;
    }
    const int f = first_parity(P);
    RETCHK(rec_half(P, A, l, f, s));
    return rec_half(P, A, l, 1 - f, s);
}

int rec_cycle(jsv_plan* P, Adj& A, const jsv_config& c, int l, bool fcyc, bool skip_pre, hipStream_t s)
{
    Hier<double>& H = P->h64;
    if (l == P->levels - 2) {
        for (int i = 0; i < c.nu1 + c.nu2; ++i) RETCHK(rec_smooth(P, A, l, s));
        return 0;
    }
    if (!skip_pre)
        for (int i = 0; i < c.nu1; ++i) RETCHK(rec_smooth(P, A, l, s));
    RETCHK(snap(P, A, l, s));
    RETCHK(restrict_(P, H, l, s));
    RETCHK(zero(H.phi[l + 1], H.g[l + 1].n, s));
    if (fcyc)
        for (int i = 0; i < c.gamma; ++i) RETCHK(rec_cycle(P, A, c, l + 1, true, false, s));
    for (int i = 0; i < c.gamma; ++i) RETCHK(rec_cycle(P, A, c, l + 1, false, false, s));
    RETCHK(prolong(P, H, l, c.nu2, s));
    for (int i = 0; i < c.nu2; ++i) RETCHK(rec_smooth(P, A, l, s));
    return 0;
}

double* pop(Adj& A, int l)
{
    auto e = A.tape.back();
    A.tape.pop_back();
    return e.first == l ? e.second : nullptr;
}

int adj_half(jsv_plan* P, Adj& A, int l, int par, const double* S, hipStream_t s)
{
    Hier<double>& H = P->h64;
    const Geo& g = H.g[l];
    dim3 gr = grid_points(g, P->dim, g.mx / 2);
    hipLaunchKernelGGL(jsv::k_adj_half_a, gr, dim3(jsv::BX, jsv::BY), 0, s, A.g[l], A.t[l], S,
                       H.cc[l], A.rbar[l], A.bx[l], A.by[l], P->dim == 3 ? A.bz[l] : nullptr,
                       A.bc[l], g, par, P->dim);
    hipLaunchKernelGGL(jsv::k_adj_half_b, gr, dim3(jsv::BX, jsv::BY), 0, s, A.g[l], A.t[l], H.cx[l],
                       H.cy[l], P->dim == 3 ? H.cz[l] : nullptr, g, 1 - par, P->dim);
    return 0;
}

int adj_smooth(jsv_plan* P, Adj& A, int l, hipStream_t s)
{
    if (P->dim == 2 && P->offdiag) {
        // the four sub-colours in reverse, each with the state after it
        Hier<double>& H = P->h64;
        const Geo& g = H.g[l];
        const dim3 ga((unsigned)((g.mx / 2 + jsv::BX - 1) / jsv::BX), (unsigned)((g.my / 2 + jsv::BY - 1) / jsv::BY), 1);
        const dim3 gb = grid_points(g, 2, g.mx - 1);
        for (int k = 3; k >= 0; --k) {
            const double* S = pop(A, l);
            if (S == nullptr) return fail("adjoint tape mismatch at level %d", l);
            const int lx = kDiagOrder[k][0], ly = kDiagOrder[k][1];
            hipLaunchKernelGGL(jsv::k_adj_diag_a, ga, dim3(jsv::BX, jsv::BY), 0, s, A.g[l], A.t[l], S, H.cc[l],
                               A.rbar[l], A.bx[l], A.by[l], A.bc[l], A.brr[l], A.blr[l], g, lx, ly);
            hipLaunchKernelGGL(jsv::k_adj_diag_b, gb, dim3(jsv::BX, jsv::BY), 0, s, A.g[l], A.t[l], H.cx[l],
                               H.cy[l], H.drr[l], H.dlr[l], g, lx, ly);
        }
        return 0;
    }
    const int f = first_parity(P);
    double* sb = pop(A, l);
    double* sa = pop(A, l);
    if (sa == nullptr || sb == nullptr) return fail("adjoint tape mismatch at level %d", l);
    RETCHK(adj_half(P, A, l, 1 - f, sb, s));
    return adj_half(P, A, l, f, sa, s);
}

int adj_cycle(jsv_plan* P, Adj& A, const jsv_config& c, int l, bool fcyc, bool skip_pre, hipStream_t s)
{
    Hier<double>& H = P->h64;
    if (l == P->levels - 2) {
        for (int i = 0; i < c.nu1 + c.nu2; ++i) RETCHK(adj_smooth(P, A, l, s));
        return 0;
    }
    for (int i = 0; i < c.nu2; ++i) RETCHK(adj_smooth(P, A, l, s));
    const Geo& gf = H.g[l];
    const Geo& gc = H.g[l + 1];
    {
        const int full = (c.nu2 <= 0 || (P->dim == 2 && P->offdiag)) ? 1 : 0;
        dim3 gr = grid_points(gc, P->dim, gc.mx - 1);
        hipLaunchKernelGGL(jsv::k_adj_prolong, gr, dim3(jsv::BX, jsv::BY), 0, s, A.g[l + 1], A.g[l],
                           gf, gc, P->dim, full);
    }
    RETCHK(zero(A.rbar[l + 1], gc.n, s));
    for (int i = 0; i < c.gamma; ++i) RETCHK(adj_cycle(P, A, c, l + 1, false, false, s));
    if (fcyc)
        for (int i = 0; i < c.gamma; ++i) RETCHK(adj_cycle(P, A, c, l + 1, true, false, s));
    // transposed restriction with the state saved before it
    double* S = pop(A, l);
    if (S == nullptr) return fail("adjoint tape mismatch (restrict) at level %d", l);
    {
        const int f = first_parity(P);
        dim3 gr = grid_points(gf, P->dim, gf.mx - 1);
        hipLaunchKernelGGL(jsv::k_adj_restrict_rbar, gr, dim3(jsv::BX, jsv::BY), 0, s, A.rb[l],
                           A.rbar[l + 1], gf, gc, P->dim, f);
        const bool od = P->dim == 2 && P->offdiag;
        hipLaunchKernelGGL(jsv::k_adj_restrict_apply, gr, dim3(jsv::BX, jsv::BY), 0, s, A.g[l], A.rb[l],
                           S, H.cx[l], H.cy[l], P->dim == 3 ? H.cz[l] : nullptr, H.cc[l], A.rbar[l],
                           A.bx[l], A.by[l], P->dim == 3 ? A.bz[l] : nullptr, A.bc[l], gf, P->dim, f,
                           od ? H.drr[l] : nullptr, od ? H.dlr[l] : nullptr, od ? A.brr[l] : nullptr,
                           od ? A.blr[l] : nullptr);
        if (od)
            hipLaunchKernelGGL(jsv::k_adj_restrict_diag, gr, dim3(jsv::BX, jsv::BY), 0, s, A.rb[l], S, A.brr[l],
                               A.blr[l], gf, f);
    }
    if (!skip_pre)
        for (int i = 0; i < c.nu1; ++i) RETCHK(adj_smooth(P, A, l, s));
    return 0;
}

int zero_rim_any(const Geo& g, int dim, double* a, hipStream_t s)
{
    for (int ph = 0; ph < dim; ++ph)
        hipLaunchKernelGGL(jsv::k_bcs, grid_exact(g.n), dim3(256), 0, s, a, g, dim, 1, -1, nullptr,
                           nullptr, nullptr, nullptr, ph);
    return 0;
}

Adj& adj_of(jsv_plan* P);

}  // namespace

// per-plan adjoint state, created on first use
#include <map>
namespace {
std::map<const jsv_plan*, Adj> g_adj;
Adj& adj_of(jsv_plan* P) { return g_adj[P]; }

void adj_drop_graphs(const jsv_plan* P)
{
    auto it = g_adj.find(P);
    if (it == g_adj.end()) return;
    for (auto& gr : it->second.graphs)
        if (gr.exec != nullptr) (void)hipGraphExecDestroy(gr.exec);
    it->second.graphs.clear();
}

// One reversed cycle: phi of level 0 holds the cycle's input.  Replayed from
// a captured graph (see run_cycle) unless graphs are off.
int adj_run_cycle(jsv_plan* P, Adj& A, const jsv_config& c, hipStream_t s)
{
    const int L = P->levels;
    auto eager = [&](hipStream_t st) -> int {
        A.used.assign(L, 0);
        A.tape.clear();
        RETCHK(rec_cycle(P, A, c, 0, c.fcycle != 0, c.skip_pre != 0, st));
        RETCHK(adj_cycle(P, A, c, 0, c.fcycle != 0, c.skip_pre != 0, st));
        if (!A.tape.empty()) return fail("adjoint tape not consumed (%zu left)", A.tape.size());
        return zero_rim_any(P->h64.g[0], P->dim, A.g[0], st);   // cycle input passes do_BCs
    };
    if (!P->use_graphs) return eager(s);
    for (auto& gr : A.graphs)
        if (gr.gamma == c.gamma && gr.fcycle == (c.fcycle != 0) && gr.skip_pre == (c.skip_pre != 0) &&
            gr.nu1 == c.nu1 && gr.nu2 == c.nu2) {
            HIPCHK(hipGraphLaunch(gr.exec, s));
            return 0;
        }
    if (P->cap_stream == nullptr) HIPCHK(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, s));
    HIPCHK(hipStreamWaitEvent(P->cap_stream, ev, 0));
    (void)hipEventDestroy(ev);
    HIPCHK(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = eager(P->cap_stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(P->cap_stream, &graph);
    if (rc != 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail("adjoint graph capture failed: %s", hipGetErrorString(e));
    Adj::Graph gr{c.gamma, c.fcycle != 0 ? 1 : 0, c.skip_pre != 0 ? 1 : 0, c.nu1, c.nu2, nullptr};
    const hipError_t ei = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail("adjoint graph instantiate failed: %s", hipGetErrorString(ei));
    A.graphs.push_back(gr);
    HIPCHK(hipGraphLaunch(gr.exec, s));
    return 0;
}

void adj_forget(const jsv_plan* P)
{
    adj_drop_graphs(P);
    g_adj.erase(P);
}
}  // namespace

extern "C" {

int jsv_forward_checkpoints(jsv_plan* P, const double* phi0, const double* rhs, const jsv_config* c,
                            double* ckpt, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (phi0 == nullptr || rhs == nullptr || ckpt == nullptr) return fail("NULL argument");
    if (P->batch != 1) return fail("checkpointed solves need an unbatched plan");
    RETCHK(check_cfg(c));
    hipStream_t s = (hipStream_t)stream;
    Hier<double>& H = P->h64;
    const long n = H.g[0].n;
    HIPCHK(hipMemcpyAsync(H.rhs[0], rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(H.phi[0], phi0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RETCHK(zero_rim_level0(P, H.phi[0], s));
    if (c->skip_pre)
        for (int i = 0; i < c->nu1; ++i) RETCHK(smooth(P, H, 0, s));
    HIPCHK(hipMemcpyAsync(ckpt, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    // the coarse program stages this operator's coarse coefficients (a program
    // built by an earlier solve holds the operator of that solve)
    RETCHK(prepare_coarse(P, H, *c, s));
    for (int64_t k = 0; k < c->max_iters; ++k) {
        RETCHK(zero_rim_level0(P, H.phi[0], s));
        RETCHK(mg_cycle(P, H, *c, 0, c->fcycle != 0, c->skip_pre != 0, s));
        HIPCHK(hipMemcpyAsync(ckpt + (k + 1) * n, H.phi[0], sizeof(double) * n,
                              hipMemcpyDeviceToDevice, s));
    }
    LAUNCHCHK();
    return 0;
}

int jsv_adjoint_full(jsv_plan* P, const double* phi0, const double* rhs, const double* ckpt,
                     const jsv_config* c, const double* g_out, double* g_phi0, double* g_rhs,
                     double* g_lam, double* g_dxx, double* g_dyy, double* g_dzz, double* g_dxy,
                     double* g_scalar, void* stream)
{
    if (P == nullptr || !P->op_set) return fail("plan missing or operator not set");
    if (phi0 == nullptr || rhs == nullptr || ckpt == nullptr || g_out == nullptr)
        return fail("NULL argument");
    if (P->batch != 1) return fail("the reverse pass needs an unbatched plan");
    if (g_lam != nullptr && !P->lam_array)
        return fail("lambda is a scalar on this plan: its cotangent goes to g_scalar[0]");
    if ((g_dxx != nullptr || g_dyy != nullptr || g_dzz != nullptr) && !P->spatial)
        return fail("D is scalar on this plan: its cotangents go to g_scalar[1..3]");
    if (g_dxy != nullptr && !(P->dim == 2 && P->offdiag)) return fail("the plan's operator has no D_xy");
    RETCHK(check_cfg(c));
    hipStream_t s = (hipStream_t)stream;
    Hier<double>& H = P->h64;
    Adj& A = adj_of(P);
    RETCHK(adj_alloc(P, A, *c));
    const int L = P->levels;
    const long n = H.g[0].n;
    for (int l = 0; l < L; ++l) {
        const long nl = H.g[l].n;
        RETCHK(zero(A.rbar[l], nl, s));
        RETCHK(zero(A.bx[l], nl, s));
        RETCHK(zero(A.by[l], nl, s));
        if (P->dim == 3) RETCHK(zero(A.bz[l], nl, s));
        RETCHK(zero(A.bc[l], nl, s));
        RETCHK(zero(A.t[l], nl, s));
        RETCHK(zero(A.rb[l], nl, s));
        RETCHK(zero(A.g[l], nl, s));
        if (!A.brr.empty()) {
            RETCHK(zero(A.brr[l], nl, s));
            RETCHK(zero(A.blr[l], nl, s));
        }
    }
    HIPCHK(hipMemcpyAsync(H.rhs[0], rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(A.g[0], g_out, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RETCHK(zero_rim_any(H.g[0], P->dim, A.g[0], s));
    for (int64_t k = c->max_iters; k >= 1; --k) {
        // recompute cycle k from its input with a tape, then reverse it
        HIPCHK(hipMemcpyAsync(H.phi[0], ckpt + (k - 1) * n, sizeof(double) * n,
                              hipMemcpyDeviceToDevice, s));
        RETCHK(adj_run_cycle(P, A, *c, s));
        LAUNCHCHK();
    }
    if (c->skip_pre && c->nu1 > 0) {
        A.used.assign(L, 0);
        A.tape.clear();
        HIPCHK(hipMemcpyAsync(H.phi[0], phi0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        RETCHK(zero_rim_level0(P, H.phi[0], s));
        for (int i = 0; i < c->nu1; ++i) RETCHK(rec_smooth(P, A, 0, s));
        for (int i = 0; i < c->nu1; ++i) RETCHK(adj_smooth(P, A, 0, s));
        RETCHK(zero_rim_any(H.g[0], P->dim, A.g[0], s));
    }
    if (g_phi0) HIPCHK(hipMemcpyAsync(g_phi0, A.g[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (g_rhs) HIPCHK(hipMemcpyAsync(g_rhs, A.rbar[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    // ---- coefficient chain ------------------------------------------------
    const Geo& g0 = H.g[0];
    // scalar parts: lambda enters every level's diagonal; a scalar D_a every
    // level's a-faces (AMat_x = D_xx idel_x^2, AMat_cc = 2 (AMat_x + AMat_y) +
    // lambda, solver_2d.py:417-421): its cotangent is the sum of each level's
    // total face bar (face bar plus the bars of the two diagonals it enters)
    if (g_scalar != nullptr) {
        RETCHK(zero(A.sums, 4, s));
        if (!P->lam_array)
            for (int l = 0; l < L; ++l)
                hipLaunchKernelGGL(jsv::k_sum_array, dim3(1), dim3(1024), 0, s, A.bc[l], H.g[l].n, A.sums, 1);
        if (!P->spatial) {
            std::vector<double*> bface[3] = {A.bx, A.by, A.bz};
            for (int a = 0; a < P->dim; ++a)
                for (int l = 0; l < L; ++l) {
                    const double id = P->idel[a] * std::ldexp(1.0, -l);
                    const Geo& g = H.g[l];
                    const long step = a == 0 ? 1 : a == 1 ? g.sy : g.sz;
                    hipLaunchKernelGGL(jsv::k_adj_face_total, grid_exact(g.n), dim3(256), 0, s, A.t[l],
                                       bface[a][l], A.bc[l], g, step, id * id, 0);
                    hipLaunchKernelGGL(jsv::k_sum_array, dim3(1), dim3(1024), 0, s, A.t[l], g.n, A.sums + 1 + a, 1);
                }
        }
        HIPCHK(hipMemcpyAsync(g_scalar, A.sums, sizeof(double) * 4, hipMemcpyDeviceToDevice, s));
    }
    if (g_dxy != nullptr) {
        // D_xy is injected (not averaged) level to level, like lambda
        RETCHK(zero(g_dxy, n, s));
        for (int l = 0; l < L; ++l) {
            const double idx = P->idel[0] * std::ldexp(1.0, -l), idy = P->idel[1] * std::ldexp(1.0, -l);
            const Geo& g = H.g[l];
            hipLaunchKernelGGL(jsv::k_adj_dxy_level, grid_exact(g.n), dim3(256), 0, s, A.chain[l], A.brr[l],
                               A.blr[l], g, idx * idy * 0.25);
            hipLaunchKernelGGL(jsv::k_adj_inject_to_fine, grid_exact(g.n), dim3(256), 0, s, g_dxy, A.chain[l], g,
                               g0, 1 << l, 2);
        }
    }
    if (g_lam) {
        RETCHK(zero(g_lam, n, s));
        for (int l = 0; l < L; ++l)
            hipLaunchKernelGGL(jsv::k_adj_inject_to_fine, grid_exact(H.g[l].n), dim3(256), 0, s, g_lam,
                               A.bc[l], H.g[l], g0, 1 << l, P->dim);
    }
    double* gD[3] = {g_dxx, g_dyy, g_dzz};
    std::vector<double*> bface[3] = {A.bx, A.by, A.bz};
    for (int a = 0; a < P->dim; ++a) {
        if (gD[a] == nullptr) continue;
        RETCHK(zero(gD[a], n, s));
        for (int l = 0; l < L; ++l) {
            const double id = P->idel[a] * std::ldexp(1.0, -l);
            const double sq = id * id;
            const Geo& g = H.g[l];
            const long step = a == 0 ? 1 : a == 1 ? g.sy : g.sz;
            hipLaunchKernelGGL(jsv::k_adj_face_total, grid_exact(g.n), dim3(256), 0, s, A.t[l],
                               bface[a][l], A.bc[l], g, step, sq, 0);
        }
        if (P->dim == 2) {
            for (int l = 0; l < L; ++l) {
                const Geo& g = H.g[l];
                hipLaunchKernelGGL(jsv::k_adj_average, grid_exact(g.n), dim3(256), 0, s, A.chain[l],
                                   A.t[l], g, 2, a);
                hipLaunchKernelGGL(jsv::k_adj_inject_to_fine, grid_exact(g.n), dim3(256), 0, s, gD[a],
                                   A.chain[l], g, g0, 1 << l, 2);
            }
        } else {
            // A.t[l] holds the bar of level l's face-averaged D; walk from the
            // coarsest level down: bar(avg_l) -> raw_l -> avg_{l-1}
            for (int l = L - 1; l >= 1; --l) {
                const Geo& g = H.g[l];
                hipLaunchKernelGGL(jsv::k_adj_average, grid_exact(g.n), dim3(256), 0, s, A.chain[l],
                                   A.t[l], g, 3, a);
                hipLaunchKernelGGL(jsv::k_adj_inject_one, grid_exact(g.n), dim3(256), 0, s,
                                   A.t[l - 1], A.chain[l], g, H.g[l - 1]);
            }
            HIPCHK(hipMemcpyAsync(gD[a], A.t[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
    }
    LAUNCHCHK();
    return 0;
}

int jsv_adjoint(jsv_plan* P, const double* phi0, const double* rhs, const double* ckpt,
                const jsv_config* c, const double* g_out, double* g_phi0, double* g_rhs,
                double* g_lam, double* g_dxx, double* g_dyy, double* g_dzz, void* stream)
{
    return jsv_adjoint_full(P, phi0, rhs, ckpt, c, g_out, g_phi0, g_rhs, g_lam, g_dxx, g_dyy, g_dzz, nullptr,
                            nullptr, stream);
}

int jsv_adjoint_device_bytes(const jsv_plan* P, int64_t* bytes)
{
    if (P == nullptr || bytes == nullptr) return fail("NULL argument");
    auto it = g_adj.find(P);
    *bytes = it == g_adj.end() ? 0 : it->second.bytes;
    return 0;
}

}  // extern "C"

// ===========================================================================
// Tangent (forward-mode) pass (FP64): the primal cycle and its tangent in
// lockstep, one launch per operation on both
// ===========================================================================
#include "jsv_tangent.hpp"

namespace {

jsv::TanCoef tan_coef(const jsv_plan* T, int l)
{
    const Hier<double>& G = T->h64;
    return {G.cx[l], G.cy[l], T->dim == 3 ? G.cz[l] : nullptr, G.cc[l],
            T->offdiag ? G.drr[l] : nullptr, T->offdiag ? G.dlr[l] : nullptr};
}

int tan_smooth(jsv_plan* P, jsv_plan* T, int l, hipStream_t s)
{
    Hier<double>& H = P->h64;
    Hier<double>& G = T->h64;
    const Geo& g = H.g[l];
    const jsv::TanCoef tc = tan_coef(T, l);
    if (P->dim == 2 && P->offdiag) {
        // the reference's sub-colour order (0,0),(1,1),(0,1),(1,0)
        static const int order[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
        const dim3 gr((unsigned)((g.mx / 2 + jsv::BX - 1) / jsv::BX),
                      (unsigned)((g.my / 2 + jsv::BY - 1) / jsv::BY), 1);
        for (int k = 0; k < 4; ++k)
            hipLaunchKernelGGL(jsv::k_rb2diag_tan, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], G.phi[l],
                               H.rhs[l], G.rhs[l], H.cx[l], H.cy[l], H.cc[l], H.drr[l], H.dlr[l], tc, g,
                               order[k][0], order[k][1]);
        return 0;
    }
    const dim3 gr = grid_points(g, P->dim, g.mx / 2);
    const int f = P->dim == 3 ? 1 : 0;   // first half: x+y+z odd (3D), x+y even (2D)
    for (int par : {f, 1 - f})
        hipLaunchKernelGGL(jsv::k_rb_tan, gr, dim3(jsv::BX, jsv::BY), 0, s, H.phi[l], G.phi[l], H.rhs[l],
                           G.rhs[l], H.cx[l], H.cy[l], P->dim == 3 ? H.cz[l] : nullptr, H.cc[l], tc, g,
                           par, P->dim);
    return 0;
}

// multigrid_cycle (solver_2d.py:582-614 / solver_3d.py:588-620) with its tangent
int tan_cycle(jsv_plan* P, jsv_plan* T, const jsv_config& c, int l, bool fcyc, bool skip_pre,
              hipStream_t s)
{
    Hier<double>& H = P->h64;
    Hier<double>& G = T->h64;
    if (l == P->levels - 2) {
        for (int i = 0; i < c.nu1 + c.nu2; ++i) RETCHK(tan_smooth(P, T, l, s));
        return 0;
    }
    if (!skip_pre)
        for (int i = 0; i < c.nu1; ++i) RETCHK(tan_smooth(P, T, l, s));
    const Geo& gf = H.g[l];
    const Geo& gc = H.g[l + 1];
    {
        // the tangent restriction reads the primal phi of level l, which the
        // primal restriction leaves alone
        const dim3 gr = grid_points(gc, P->dim, gc.mx - 1);
        const bool od = P->dim == 2 && P->offdiag;
        hipLaunchKernelGGL(jsv::k_restrict_tan, gr, dim3(jsv::BX, jsv::BY), 0, s, G.rhs[l + 1], H.phi[l],
                           G.phi[l], G.rhs[l], H.cx[l], H.cy[l], P->dim == 3 ? H.cz[l] : nullptr, H.cc[l],
                           od ? H.drr[l] : nullptr, od ? H.dlr[l] : nullptr, tan_coef(T, l), gf, gc, P->dim,
                           G.phi[l + 1]);
    }
    RETCHK(restrict_(P, H, l, s, H.phi[l + 1]));
    if (fcyc)
        for (int i = 0; i < c.gamma; ++i) RETCHK(tan_cycle(P, T, c, l + 1, true, false, s));
    for (int i = 0; i < c.gamma; ++i) RETCHK(tan_cycle(P, T, c, l + 1, false, false, s));
    RETCHK(prolong(P, H, l, c.nu2, s));
    RETCHK(prolong(T, G, l, c.nu2, s));
    for (int i = 0; i < c.nu2; ++i) RETCHK(tan_smooth(P, T, l, s));
    return 0;
}

}  // namespace

extern "C" {

int jsv_tangent(jsv_plan* P, jsv_plan* T, const double* phi0, const double* rhs, const double* phi0_dot,
                const double* rhs_dot, const jsv_config* c, double* phi_out, double* phi_dot_out,
                void* stream)
{
    if (P == nullptr || T == nullptr || !P->op_set || !T->op_set)
        return fail("plan missing or operator not set");
    if (phi0 == nullptr || rhs == nullptr || phi_dot_out == nullptr) return fail("NULL argument");
    if (P == T) return fail("the tangent needs its own plan");
    if (P->dim != T->dim || P->levels != T->levels || P->m[0] != T->m[0] || P->m[1] != T->m[1] ||
        P->m[2] != T->m[2])
        return fail("primal and tangent plans differ in shape");
    if (P->batch != 1 || T->batch != 1) return fail("the tangent pass needs unbatched plans");
    if (P->offdiag != T->offdiag) return fail("primal and tangent plans differ in D_xy");
    RETCHK(check_cfg(c));
    hipStream_t s = (hipStream_t)stream;
    Hier<double>& H = P->h64;
    Hier<double>& G = T->h64;
    const long n = H.g[0].n;
    HIPCHK(hipMemcpyAsync(H.rhs[0], rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(H.phi[0], phi0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (rhs_dot) HIPCHK(hipMemcpyAsync(G.rhs[0], rhs_dot, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    else RETCHK(zero(G.rhs[0], n, s));
    if (phi0_dot) HIPCHK(hipMemcpyAsync(G.phi[0], phi0_dot, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    else RETCHK(zero(G.phi[0], n, s));
    RETCHK(zero_rim_level0(P, H.phi[0], s));
    RETCHK(zero_rim_level0(P, G.phi[0], s));
    if (c->skip_pre)
        for (int i = 0; i < c->nu1; ++i) RETCHK(tan_smooth(P, T, 0, s));
    for (int64_t k = 0; k < c->max_iters; ++k) {
        RETCHK(zero_rim_level0(P, H.phi[0], s));
        RETCHK(zero_rim_level0(P, G.phi[0], s));
        RETCHK(tan_cycle(P, T, *c, 0, c->fcycle != 0, c->skip_pre != 0, s));
        LAUNCHCHK();
    }
    if (phi_out) HIPCHK(hipMemcpyAsync(phi_out, H.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(phi_dot_out, G.phi[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    LAUNCHCHK();
    return 0;
}

}  // extern "C"
