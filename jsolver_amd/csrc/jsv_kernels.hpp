// Device kernels of the multigrid path for gfx950 (CDNA4, wave64).
//
// Layout: every level stores phi, rhs and its coefficients as ghost-padded
// row-major arrays of the same extents (nz, ny, nx) = (mz+3, my+3, mx+3); in 2D
// nz = 1.  Coefficients use the phi index of the point they belong to:
//   cx[p]  face coefficient between p and p+x (the reference's AMat_rc /
//          AMat_rcc at p; its AMat_lc / AMat_lcc at p is cx[p-1]),
//   cy[p], cz[p] likewise along y and z,
//   cc[p]  the diagonal AMat_cc / AMat_ccc,
//   drr[p] 2D off-diagonal coupling p <-> p+(1,1) (AMat_rr; AMat_ll at p is
//          drr[p-(1,1)]), dlr[p] coupling p <-> p+(1,-1) in (y, x) (AMat_lr;
//          AMat_rl at p is dlr[p+(-1,+1)]).
// Storing the shared faces once (instead of the reference's separate left and
// right arrays) is bit-identical: both are the same product D_face * idel^2.
//
// Arithmetic follows the reference's expressions term by term and is compiled
// without FMA contraction, so the FP64 kernels reproduce an IEEE evaluation of
// the reference formulas exactly (the CPU oracle does the same in NumPy).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace jsv {

struct Geo {
    int nx, ny, nz;   // padded extents; in 2D nz = number of batched problems
    int mx, my, mz;   // resolutions (mz = 0 in 2D)
    long sy, sz;      // strides in elements (2D: sz = one problem's array)
    long n;           // nx * ny * nz
};

constexpr int BX = 64;   // one wave across x
constexpr int BY = 4;    // 4 waves per workgroup

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2);
// remap the launch index so XCD k works on a contiguous run of tiles and the
// halo rows neighbouring tiles share are fetched once per L2.
__device__ __forceinline__ int xcd_tile(int id, int n)
{
    const int per = n >> 3;
    return id < (per << 3) ? (id & 7) * per + (id >> 3) : id;
}

// The launch's block index remapped the same way for the one-point-per-thread
// 3D kernels: XCD k works on the k-th contiguous eighth of the row-major
// block order (a z-slab of the level), so the neighbouring rows and planes a
// stencil reads mostly belong to blocks of the same XCD and are fetched once
// into its L2.
__device__ __forceinline__ uint3 xcd_block()
{
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int t = xcd_tile((int)(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z)), gx * gy * (int)gridDim.z);
    return make_uint3((unsigned)(t % gx), (unsigned)((t / gx) % gy), (unsigned)(t / (gx * gy)));
}

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel half sweeps (solver_2d.py:670-692, solver_3d.py:679-706).
// The reference's sub-colours split into two independent halves: in 2D the
// sub-lattices (0,0),(1,1) are the points with x+y even and (0,1),(1,0) those
// with x+y odd; in 3D the first four sub-colours are the points with x+y+z odd,
// the last four those with x+y+z even.  Points of one half never neighbour
// each other in the 5/7-point stencil, so one launch per half reproduces the
// reference's sequential sub-sweeps exactly.  Boundary nodes (index 1, m+1) and
// ghosts are never written and stay zero, which is what the reference's
// do_BCs between sub-sweeps guarantees.
// ---------------------------------------------------------------------------
// Per-point updates (shared by the launch-per-half kernels below and the
// single-workgroup coarse-level program kernel), in the reference's term order.
template <typename T>
__device__ __forceinline__ void rb2_point(T* __restrict__ phi, const T* __restrict__ rhs,
                                          const T* __restrict__ cx, const T* __restrict__ cy,
                                          const T* __restrict__ cc, long p, long sy)
{
    T acc = cy[p] * phi[p + sy] + cy[p - sy] * phi[p - sy];
    acc = acc + cx[p] * phi[p + 1];
    acc = acc + cx[p - 1] * phi[p - 1];
    acc = acc - rhs[p];
    phi[p] = acc / cc[p];
}

template <typename T>
__device__ __forceinline__ void rb2diag_point(T* __restrict__ phi, const T* __restrict__ rhs,
                                              const T* __restrict__ cx, const T* __restrict__ cy,
                                              const T* __restrict__ cc, const T* __restrict__ drr,
                                              const T* __restrict__ dlr, long p, long sy)
{
    T acc = cy[p] * phi[p + sy] + cy[p - sy] * phi[p - sy];
    acc = acc + cx[p] * phi[p + 1];
    acc = acc + cx[p - 1] * phi[p - 1];
    acc = acc - rhs[p];
    T off = drr[p - sy - 1] * phi[p - sy - 1] + dlr[p - sy + 1] * phi[p - sy + 1];
    off = off + dlr[p] * phi[p + sy - 1];
    off = off + drr[p] * phi[p + sy + 1];
    acc = acc + off;
    phi[p] = acc / cc[p];
}

template <typename T>
__device__ __forceinline__ void rb3_point(T* __restrict__ phi, const T* __restrict__ rhs,
                                          const T* __restrict__ cx, const T* __restrict__ cy,
                                          const T* __restrict__ cz, const T* __restrict__ cc, long p,
                                          long sy, long sz)
{
    T acc = cz[p] * phi[p + sz] + cz[p - sz] * phi[p - sz];
    acc = acc + cy[p] * phi[p + sy];
    acc = acc + cy[p - sy] * phi[p - sy];
    acc = acc + cx[p] * phi[p + 1];
    acc = acc + cx[p - 1] * phi[p - 1];
    acc = acc - rhs[p];
    phi[p] = acc / cc[p];
}

// 2D kernels: blockIdx.z is the problem of a batch; act (nullable) holds a
// per-problem "still iterating" flag.
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb2(T* __restrict__ phi, const T* __restrict__ rhs, const T* __restrict__ cx,
      const T* __restrict__ cy, const T* __restrict__ cc, Geo g, int par, const int* __restrict__ act)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > g.my) return;
    const int x = 2 + ((par + y) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    rb2_point(phi, rhs, cx, cy, cc, (long)b * g.sz + (long)y * g.sy + x, g.sy);
}

// 2D with D_xy: the 9-point stencil couples the sub-lattices diagonally, so
// each of the reference's four sub-colours (lx, ly) is its own launch.
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb2_diag(T* __restrict__ phi, const T* __restrict__ rhs, const T* __restrict__ cx,
           const T* __restrict__ cy, const T* __restrict__ cc, const T* __restrict__ drr,
           const T* __restrict__ dlr, Geo g, int lx, int ly, const int* __restrict__ act)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int y = (ly ? 2 : 3) + 2 * (blockIdx.y * BY + threadIdx.y);
    if (y > g.my) return;
    const int x = (lx ? 2 : 3) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    rb2diag_point(phi, rhs, cx, cy, cc, drr, dlr, (long)b * g.sz + (long)y * g.sy + x, g.sy);
}

// z0: first plane of the launch (grid.z planes from z0); the slab schedule
// launches one z-slab at a time.
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb3(T* __restrict__ phi, const T* __restrict__ rhs, const T* __restrict__ cx,
      const T* __restrict__ cy, const T* __restrict__ cz, const T* __restrict__ cc, Geo g, int par,
      int z0, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const uint3 bk = xcd_block();
    const int y = 2 + bk.y * BY + threadIdx.y;
    const int z = z0 + bk.z;
    if (y > g.my) return;
    const int x = 2 + ((par + y + z) & 1) + 2 * (bk.x * BX + threadIdx.x);
    if (x > g.mx) return;
    rb3_point(phi, rhs, cx, cy, cz, cc, (long)z * g.sz + (long)y * g.sy + x, g.sy, g.sz);
}

// One RBGS half over the whole level that also snapshots phi after the half
// (for the adjoint's tape): points of the half are updated and stored to
// both, the others copied.  Reads never touch points the half writes, so
// the in-place update has no race.  FP64 only; 2D (dim 2, one problem) or 3D.
__global__ void __launch_bounds__(256)
k_rb_snap(double* __restrict__ phi, double* __restrict__ snap, const double* __restrict__ rhs,
          const double* __restrict__ cx, const double* __restrict__ cy, const double* __restrict__ cz,
          const double* __restrict__ cc, Geo g, int par, int dim)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long n = dim == 3 ? g.n : g.sz;
    if (i >= n) return;
    const int z = dim == 3 ? (int)(i / g.sz) : 0;
    const long r = i - (long)z * g.sz;
    const int y = (int)(r / g.sy), x = (int)(r - (long)y * g.sy);
    const bool unk = x >= 2 && x <= g.mx && y >= 2 && y <= g.my && (dim == 2 || (z >= 2 && z <= g.mz));
    // the half's points are x = 2 + ((par + y [+ z]) & 1) + 2k, i.e. x + y [+ z] = par mod 2
    const int colour = dim == 3 ? ((x + y + z) & 1) : ((x + y) & 1);
    if (unk && colour == par) {
        if (dim == 3) rb3_point(phi, rhs, cx, cy, cz, cc, i, g.sy, g.sz);
        else rb2_point(phi, rhs, cx, cy, cc, i, g.sy);
    }
    snap[i] = phi[i];
}

// Prolongation corrections (solver_2d.py:649-668, solver_3d.py:654-677), shared by
// the prolongation kernels and the pair sweep that applies the correction on the fly.
// Correction at fine point (x, y[, z]) from coarse array pc (problem offset
// included); false where the point is skipped (full == 0).
// C2(J, K) reads the coarse correction at coarse point (J, K)
template <typename T, typename Coarse>
__device__ __forceinline__ bool prolong2_corr_at(Coarse C2, int x, int y, int full, T& corr)
{
    const int ox = x & 1, oy = y & 1;
    const int Xo = (x + 1) >> 1, Yo = (y + 1) >> 1, Xl = x >> 1, Yl = y >> 1;
    if (ox && oy) {
        corr = C2(Yo, Xo);
    } else if (oy) {
        corr = (T)0.5 * (C2(Yo, Xl + 1) + C2(Yo, Xl));
    } else if (ox) {
        corr = (T)0.5 * (C2(Yl + 1, Xo) + C2(Yl, Xo));
    } else {
        if (!full) return false;
        T s = C2(Yl + 1, Xl + 1) + C2(Yl, Xl + 1);
        s = s + C2(Yl + 1, Xl);
        s = s + C2(Yl, Xl);
        corr = (T)0.25 * s;
    }
    return true;
}

template <typename T>
__device__ __forceinline__ bool prolong2_corr(const T* __restrict__ pc, long csy, int x, int y,
                                              int full, T& corr)
{
    return prolong2_corr_at<T>([&](int J, int K) -> T { return pc[(long)J * csy + K]; }, x, y, full, corr);
}

// C3(I, J, K) reads the coarse correction at coarse point (I, J, K)
template <typename T, typename Coarse>
__device__ __forceinline__ bool prolong3_corr_at(Coarse C3, int x, int y, int z, int full, T& corr)
{
    const int ox = x & 1, oy = y & 1, oz = z & 1;
    const int Xo = (x + 1) >> 1, Yo = (y + 1) >> 1, Zo = (z + 1) >> 1;
    const int Xl = x >> 1, Yl = y >> 1, Zl = z >> 1;
    const int nodd = ox + oy + oz;
    if (nodd == 3) {
        corr = C3(Zo, Yo, Xo);
    } else if (nodd == 2) {
        if (!ox)      corr = (T)0.5 * (C3(Zo, Yo, Xl + 1) + C3(Zo, Yo, Xl));
        else if (!oy) corr = (T)0.5 * (C3(Zo, Yl + 1, Xo) + C3(Zo, Yl, Xo));
        else          corr = (T)0.5 * (C3(Zl + 1, Yo, Xo) + C3(Zl, Yo, Xo));
    } else if (nodd == 1) {
        if (!full) return false;
        T s;
        if (oz) {
            s = C3(Zo, Yl + 1, Xl + 1) + C3(Zo, Yl, Xl + 1);
            s = s + C3(Zo, Yl + 1, Xl);
            s = s + C3(Zo, Yl, Xl);
        } else if (oy) {
            s = C3(Zl + 1, Yo, Xl + 1) + C3(Zl, Yo, Xl + 1);
            s = s + C3(Zl + 1, Yo, Xl);
            s = s + C3(Zl, Yo, Xl);
        } else {
            s = C3(Zl + 1, Yl + 1, Xo) + C3(Zl, Yl + 1, Xo);
            s = s + C3(Zl + 1, Yl, Xo);
            s = s + C3(Zl, Yl, Xo);
        }
        corr = (T)0.25 * s;
    } else {
        T s = C3(Zl + 1, Yl + 1, Xl + 1) + C3(Zl, Yl + 1, Xl + 1);
        s = s + C3(Zl + 1, Yl, Xl + 1);
        s = s + C3(Zl, Yl, Xl + 1);
        s = s + C3(Zl + 1, Yl + 1, Xl);
        s = s + C3(Zl, Yl + 1, Xl);
        s = s + C3(Zl + 1, Yl, Xl);
        s = s + C3(Zl, Yl, Xl);
        corr = (T)0.125 * s;
    }
    return true;
}

// prolong3_corr_at at a point with x + y + z even, which has all three
// coordinates even (eight terms) or exactly one (two terms along that axis):
// the same terms in the same order, without the branches for the cases such a
// point never takes.  Along an axis the terms sit at v/2 + 1 and v/2 + (v & 1),
// which for an odd v are both (v + 1)/2.
template <typename T, typename Coarse>
__device__ __forceinline__ T prolong3_corr_even(Coarse C3, int x, int y, int z)
{
    const int Xl = x >> 1, Yl = y >> 1, Zl = z >> 1;
    if ((x | y | z) & 1) {
        return (T)0.5 * (C3(Zl + 1, Yl + 1, Xl + 1) + C3(Zl + (z & 1), Yl + (y & 1), Xl + (x & 1)));
    }
    T s = C3(Zl + 1, Yl + 1, Xl + 1) + C3(Zl, Yl + 1, Xl + 1);
    s = s + C3(Zl + 1, Yl, Xl + 1);
    s = s + C3(Zl, Yl, Xl + 1);
    s = s + C3(Zl + 1, Yl + 1, Xl);
    s = s + C3(Zl, Yl + 1, Xl);
    s = s + C3(Zl + 1, Yl, Xl);
    s = s + C3(Zl, Yl, Xl);
    return (T)0.125 * s;
}

template <typename T>
__device__ __forceinline__ bool prolong3_corr(const T* __restrict__ pc, long csy, long csz, int x,
                                              int y, int z, int full, T& corr)
{
    return prolong3_corr_at<T>([&](int I, int J, int K) -> T { return pc[(long)I * csz + (long)J * csy + K]; },
                               x, y, z, full, corr);
}

// First RBGS half of the post-smoothing with the prolongation phi += P e
// (solver_2d.py:612, 649-668; solver_3d.py:618, 654-677) applied on the fly:
// this half reads only the other half's values, each taken as stored value
// plus its correction (computed as the prolongation kernels do, unknowns
// only), and with nu2 > 0 nothing reads the other half's corrected values
// later (the second half overwrites them without reading them) nor this
// half's own old values.  Same bits as k_prolong3 / k_prolong2 then the half.
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb3_pro(T* __restrict__ phi, const T* __restrict__ rhs, const T* __restrict__ cx,
          const T* __restrict__ cy, const T* __restrict__ cz, const T* __restrict__ cc, Geo g, int par,
          const T* __restrict__ pc, Geo gc, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const uint3 bk = xcd_block();
    const int y = 2 + bk.y * BY + threadIdx.y;
    const int z = 2 + bk.z;
    if (y > g.my) return;
    const int x = 2 + ((par + y + z) & 1) + 2 * (bk.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long sy = g.sy, sz = g.sz;
    const long p = (long)z * sz + (long)y * sy + x;
    // the neighbours have x + y + z = par + 1 (mod 2): even for par == 1
    auto C3 = [&](int I, int J, int K) -> T { return pc[(long)I * gc.sz + (long)J * gc.sy + K]; };
    auto nb = [&](int xx, int yy, int zz, long q) -> T {
        T v = phi[q];
        if (xx >= 2 && xx <= g.mx && yy >= 2 && yy <= g.my && zz >= 2 && zz <= g.mz) {
            T corr;
            if (par == 1) v = v + prolong3_corr_even<T>(C3, xx, yy, zz);
            else if (prolong3_corr_at<T>(C3, xx, yy, zz, 0, corr)) v = v + corr;
        }
        return v;
    };
    T acc = cz[p] * nb(x, y, z + 1, p + sz) + cz[p - sz] * nb(x, y, z - 1, p - sz);
    acc = acc + cy[p] * nb(x, y + 1, z, p + sy);
    acc = acc + cy[p - sy] * nb(x, y - 1, z, p - sy);
    acc = acc + cx[p] * nb(x + 1, y, z, p + 1);
    acc = acc + cx[p - 1] * nb(x - 1, y, z, p - 1);
    acc = acc - rhs[p];
    phi[p] = acc / cc[p];
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb2_pro(T* __restrict__ phi, const T* __restrict__ rhs, const T* __restrict__ cx,
          const T* __restrict__ cy, const T* __restrict__ cc, Geo g, int par, const int* __restrict__ act,
          const T* __restrict__ pc, Geo gc)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > g.my) return;
    const int x = 2 + ((par + y) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long sy = g.sy;
    const long p = (long)b * g.sz + (long)y * sy + x;
    const T* pcb = pc + (long)b * gc.sz;
    auto nb = [&](int xx, int yy, long q) -> T {
        T v = phi[q];
        T corr;
        if (xx >= 2 && xx <= g.mx && yy >= 2 && yy <= g.my && prolong2_corr(pcb, gc.sy, xx, yy, 0, corr))
            v = v + corr;
        return v;
    };
    T acc = cy[p] * nb(x, y + 1, p + sy) + cy[p - sy] * nb(x, y - 1, p - sy);
    acc = acc + cx[p] * nb(x + 1, y, p + 1);
    acc = acc + cx[p - 1] * nb(x - 1, y, p - 1);
    acc = acc - rhs[p];
    phi[p] = acc / cc[p];
}

// ---------------------------------------------------------------------------
// Fused RBGS sweep (red half then black half), out of place, marching in z.
//
// One workgroup owns TY rows of the level (all of x) and a chunk of planes
// [z0, z1).  Per step it loads the black values of phi_in on plane z+1 (rows
// y0-2 .. y0+TY+1) into an LDS ring, computes the new red values of plane z on
// rows y0-1 .. y0+TY (one halo row each side, recomputed by the neighbour
// tile too, from the same inputs, so both get identical bits) into a second
// LDS ring, then the new black values of plane z-1 on its own rows from the
// red ring, and writes both colours of plane z-1 to phi_out.  phi_in is only
// read, phi_out only written, so tiles never race; each array crosses HBM
// about once per sweep instead of once per half.  Both rings store one colour
// compactly: the point (x, y, z) of that colour lives at column x >> 1.
// "Red" is the first half of the reference's sub-colour order (x+y+z odd).
// Rim values of phi_out are never written (zero from allocation).
// ---------------------------------------------------------------------------
extern __shared__ double jsv_dyn_smem[];

template <typename T, int TY, int NT>
__global__ void __launch_bounds__(NT)
k_rbsweep3(const T* __restrict__ pin, T* __restrict__ pout, const T* __restrict__ rhs,
           const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
           const T* __restrict__ cc, Geo g, int zchunk, int ytiles, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    constexpr int PR = TY + 4;                  // rows of the phi_in ring: y0-2 .. y0+TY+1
    constexpr int RR = TY + 2;                  // rows of the red ring:    y0-1 .. y0+TY
    constexpr int NHMAX = 258;                  // compact columns covered by the register prefetch (m <= 512)
    constexpr int PMAX = (PR * NHMAX + NT - 1) / NT;
    const int nh = (g.nx + 1) >> 1;             // compact columns per row
    T* P = reinterpret_cast<T*>(jsv_dyn_smem);  // [3][PR][nh]  black values of phi_in
    T* R = P + 3 * PR * nh;                     // [3][RR][nh]  new red values
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int y0 = 2 + (tile % ytiles) * TY;
    const int z0 = 2 + (tile / ytiles) * zchunk;
    int z1 = z0 + zchunk;
    if (z1 > g.mz + 1) z1 = g.mz + 1;
    const long sy = g.sy, sz = g.sz;
    const bool prefetch = nh <= NHMAX;

    for (int i = tid; i < 3 * RR * nh; i += NT) R[i] = (T)0;

    // black values (x+y+z even) of phi_in on plane z, element i of the slot
    auto fetch = [&](int z, int i) -> T {
        const int r = i / nh, k = i - r * nh;
        const int y = y0 - 2 + r;
        const int x = 2 * k + ((y + z) & 1);
        if (z >= 0 && z < g.nz && y >= 0 && y < g.ny && x < g.nx) return pin[(long)z * sz + (long)y * sy + x];
        return (T)0;
    };
    auto load_plane = [&](int z) {
        T* dst = P + (z % 3) * PR * nh;
        for (int i = tid; i < PR * nh; i += NT) dst[i] = fetch(z, i);
    };
    // next plane held in registers while the current one is computed
    T pre[PMAX];
    auto issue = [&](int z) {
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            const int i = tid + j * NT;
            pre[j] = (i < PR * nh) ? fetch(z, i) : (T)0;
        }
    };
    auto commit = [&](int z) {
        T* dst = P + (z % 3) * PR * nh;
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            const int i = tid + j * NT;
            if (i < PR * nh) dst[i] = pre[j];
        }
    };
    auto Pv = [&](int z, int y, int x) -> T {     // black phi_in at (x, y, z)
        return P[((z % 3) * PR + (y - y0 + 2)) * nh + (x >> 1)];
    };
    auto Rv = [&](int z, int y, int x) -> T {     // new red at (x, y, z)
        return R[((z % 3) * RR + (y - y0 + 1)) * nh + (x >> 1)];
    };

    load_plane(z0 - 2);
    load_plane(z0 - 1);
    if (prefetch) issue(z0);
    for (int zs = z0 - 1; zs <= z1; ++zs) {
        if (prefetch) {
            commit(zs + 1);
            if (zs + 2 <= z1 + 1) issue(zs + 2);      // in flight during this step's compute
        } else {
            load_plane(zs + 1);
        }
        __syncthreads();
        // ---- red on plane zs, rows y0-1 .. y0+TY
        {
            T* Rs = R + (zs % 3) * RR * nh;
            const bool zok = (zs >= 2 && zs <= g.mz);
#pragma unroll 4
            for (int i = tid; i < RR * nh; i += NT) {
                const int r = i / nh, k = i - r * nh;
                const int y = y0 - 1 + r;
                const int x = 2 * k + ((1 + y + zs) & 1);   // x + y + z odd
                if (!zok || y < 2 || y > g.my || x < 2 || x > g.mx) {
                    Rs[i] = (T)0;
                    continue;
                }
                const long p = (long)zs * sz + (long)y * sy + x;
                T acc = cz[p] * Pv(zs + 1, y, x) + cz[p - sz] * Pv(zs - 1, y, x);
                acc = acc + cy[p] * Pv(zs, y + 1, x);
                acc = acc + cy[p - sy] * Pv(zs, y - 1, x);
                acc = acc + cx[p] * Pv(zs, y, x + 1);
                acc = acc + cx[p - 1] * Pv(zs, y, x - 1);
                acc = acc - rhs[p];
                Rs[i] = acc / cc[p];
            }
        }
        __syncthreads();
        // ---- black on plane zb = zs-1, rows y0 .. y0+TY-1; write both colours
        const int zb = zs - 1;
        if (zb >= z0 && zb < z1) {
#pragma unroll 4
            for (int i = tid; i < TY * g.nx; i += NT) {
                const int r = i / g.nx, x = i - r * g.nx;
                const int y = y0 + r;
                if (y > g.my || x < 2 || x > g.mx) continue;
                const long q = (long)zb * sz + (long)y * sy + x;
                if (((x + y + zb) & 1) == 1) {          // red: already computed
                    pout[q] = Rv(zb, y, x);
                    continue;
                }
                T acc = cz[q] * Rv(zb + 1, y, x) + cz[q - sz] * Rv(zb - 1, y, x);
                acc = acc + cy[q] * Rv(zb, y + 1, x);
                acc = acc + cy[q - sy] * Rv(zb, y - 1, x);
                acc = acc + cx[q] * Rv(zb, y, x + 1);
                acc = acc + cx[q - 1] * Rv(zb, y, x - 1);
                acc = acc - rhs[q];
                pout[q] = acc / cc[q];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Pair-per-thread variant of the fused sweep, with the coefficient stream
// software-pipelined one plane ahead (NP consecutive pairs per thread).
//
// Same tiles, rings and step order as k_rbsweep3, but each thread owns one
// column pair (x = 2k, 2k+1) of one ring row for the whole z-march.  In every
// plane one point of a pair is red and the other black, and the black point
// of plane z-1 sits at the x of the red point of plane z.  So per step a
// thread loads the coefficients of BOTH points of plane zs+1 (13 loads,
// every element of the cache lines used: cz of the plane below and cx of the
// left neighbour are carried or shared; pairs load as one vector), uses
// plane zs's red half for the red
// update now and keeps plane zs's black half in registers for the black
// update one step later; its own red value of plane zs-1 is written out from
// a register.  The two coefficient sets alternate between two register
// blocks (the loop is unrolled by two), so the loads issued for plane zs+1
// stay in flight across the step's barriers and compute; the black half
// waits in LDS.  With PRO the prolongation that precedes a post-smoothing
// sweep is applied on the fly to the black phi_in values (the only ones the
// sweep reads).  Requires (TY + 2) * (mx / 2) <= NT, (nx + 1) / 2 <= NHMAX
// and the level below 4 GiB; rim values of phi_in are zero.
// ---------------------------------------------------------------------------
// First-half residuals of a 3D level are stored compactly: point (z, y, x)
// with x + y + z odd, z, y, x in 2..m, at ((z-2)(my-1) + y-2) * mx/2 + (x-2)/2
// (x = 2 + parity + 2i, so (x-2)/2 = i): half the bytes of a full array.
__device__ __forceinline__ long res3_at(const Geo& g, int z, int y, int x)
{
    return ((long)(z - 2) * (g.my - 1) + (y - 2)) * (g.mx >> 1) + ((x - 2) >> 1);
}

// sums / maxima over the 64 lanes of a wave (every lane gets the result)
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// N consecutive elements, loaded as one dword-aligned vector access
template <typename T, int N>
struct __attribute__((aligned(sizeof(T)))) PairVec {
    T v[N];
};

// which element of a column pair is red on a plane: a compile-time constant
// (PairElem) or computed from y + z (PairElemRT)
template <int V>
struct PairElem {
    __device__ static constexpr int at(int, int) { return V; }
};
struct PairElemRT {
    __device__ static int at(int y, int z) { return ((y + z) & 1) != 0 ? 0 : 1; }
};

// the coefficients of NP consecutive column pairs of one row and plane
template <typename T, int NP>
struct PairGrp {
    T cxm;                     // cx at x = 2k - 1 (left of the first pair)
    PairVec<T, 2 * NP> cx, cy, cym, cz, rh, dg;   // at x = 2k .. 2k + 2NP - 1; cym: row y - 1
};

template <typename T, int TY, int NT, int NHMAX, bool PRO = false, int NP = 1, int WPE = 1, bool WR = false,
          bool RES = false, bool FR = false>
__global__ void __launch_bounds__(NT, WPE)
k_rbsweep3p(const T* __restrict__ pin, T* __restrict__ pout, const T* __restrict__ rhs,
            const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
            const T* __restrict__ cc, Geo g, int zchunk, int ytiles, const T* __restrict__ pc, Geo gc,
            const int* __restrict__ act, double* __restrict__ res_part, int dn_in, T* __restrict__ rfo)
{
    if (act != nullptr && act[0] == 0) return;
    static_assert(!(RES && PRO), "the residual is of phi_in as given");
    static_assert(!(FR && (PRO || RES)), "the first-half residuals come from the plain sweep");
    constexpr int PR = TY + 4;                  // rows of the phi_in ring: y0-2 .. y0+TY+1
    constexpr int RR = TY + 2;                  // rows of the red ring / pair rows: y0-1 .. y0+TY
    constexpr int PMAX = (PR * NHMAX + NT - 1) / NT;   // ring elements per thread (nh <= NHMAX)
    constexpr int CRR = TY / 2 + 4;             // coarse rows under the phi_in ring rows
    constexpr int CMAX = (CRR * (NHMAX + 1) + NT - 1) / NT;   // coarse ring elements per thread
    // ring rows are NHMAX wide (the host sizes LDS for that): constant strides
    // fold the neighbour offsets into the LDS instructions; columns at or past
    // (nx + 1) / 2 hold zeros and are never read
    constexpr int nh = NHMAX;
    const int nhe = g.mx >> 1;                  // pairs per row
    const int gpr = nhe / NP;                   // thread columns per row (NP pairs each)
    T* P = reinterpret_cast<T*>(jsv_dyn_smem);  // [3][PR][nh]
    T* R = P + 3 * PR * nh;                     // [3][RR][nh]
    T* Hb = R + 3 * RR * nh;                    // [NP][8][NT] this thread's held black halves
    T* Cr = Hb + 8 * NP * NT;                   // PRO: [3][CRR][gc.nx] coarse planes of the correction
    T* Ro = Cr + (PRO ? 3 * CRR * (NHMAX + 1) : 0);   // RES: [3][RR][nh] red values of phi_in
    T* Pb = Ro + 3 * RR * NHMAX;                // RES: [NP][NT] this thread's held black values of phi_in
    // FR: for plane zs - 1 and each red row y0 + 1 .. y0 + TY - 2 (row i =
    // y - y0 - 1), the black values of that row (the residual's x
    // neighbours) and the products of the black points of the rows below and
    // above with their faces toward it (its y terms); per thread, the red rhs
    // of the last two planes and one partial residual per pair
    T* Bk = Ro;                                 // FR: [TY - 2][nh] phi of row y
    T* Bu = Bk + (TY - 2) * NHMAX;              // FR: [TY - 2][nh] cy(y - 1) * phi(y - 1)
    T* Bd = Bu + (TY - 2) * NHMAX;              // FR: [TY - 2][nh] cy(y) * phi(y + 1)
    T* Hf = Bd + (TY - 2) * NHMAX;              // FR: [4][NP][NT]
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int y0 = 2 + (tile % ytiles) * TY;
    const int zci = tile / ytiles;              // z-chunk index
    // dn (below): this workgroup marches downward.  dn_in bit 0: the
    // direction; bit 1: chunks alternate, chunk zci marching in direction
    // bit 0 ^ (zci & 1), and chunk bounds are those of the upward march in
    // stored planes whatever the direction, so neighbouring chunks reach
    // their shared boundary planes at the same time (both at the start or
    // both at the end of the march)
    const bool zalt = (dn_in & 2) != 0;
    const int dn = RES ? 0 : ((dn_in & 1) ^ (zalt ? (zci & 1) : 0));
    int z0 = 2 + zci * zchunk;
    int z1 = z0 + zchunk;
    if (z1 > g.mz + 1) z1 = g.mz + 1;
    if (zalt && dn) {
        // stored planes z0 .. z1 - 1 are march planes mz + 3 - z1 .. mz + 2 - z0
        const int u0 = g.mz + 3 - z1;
        z1 = g.mz + 3 - z0;
        z0 = u0;
    }
    const long sy = g.sy, sz = g.sz;

    for (int i = tid; i < 3 * RR * nh; i += NT) R[i] = (T)0;
    if constexpr (RES)
        for (int i = tid; i < 3 * RR * nh; i += NT) Ro[i] = (T)0;
    if constexpr (FR)
        for (int i = tid; i < 3 * (TY - 2) * nh; i += NT) Bk[i] = (T)0;

    const bool has = tid < RR * gpr;
    const int r = has ? tid / gpr : 0;
    const int k = has ? 1 + NP * (tid - r * gpr) : 1;   // first pair of this thread
    const int y = y0 - 1 + r;
    const bool ld = has && y < g.ny;
    const bool unk = has && y >= 2 && y <= g.my;
    const bool inner = unk && r >= 1 && r <= TY;
    // threads without an in-bounds pair load row 1 (values unused), so every
    // thread issues the same loads.  Byte offsets are 32-bit (levels below
    // 4 GiB, checked by the host): one offset register serves all arrays,
    // whose bases stay in scalar registers.
    const unsigned rowb = (unsigned)(((ld ? y : 1) * g.sy + 2 * k) * (long)sizeof(T));
    const unsigned syb = (unsigned)(g.sy * sizeof(T)), szb = (unsigned)(g.sz * sizeof(T));
    // dn: the chunks are marched downward, so that a sweep can visit the
    // planes in the reverse order of the pass before it (the planes that pass
    // touched last are still in the Infinity Cache).  The kernel works in march
    // coordinates u; plane u is stored at pl(u) = nz - 1 - u.  The mirror maps
    // unknowns to unknowns and keeps every parity (mz + 2 is even); the z face
    // above march plane u is the stored cz of plane pl(u + 1) (clamped for the
    // prefetch past the last plane, whose values are never used).  The two z
    // terms of a stencil sum swap places, and that first sum is commutative,
    // so the bits are those of the upward march.  The variant with the
    // residual sum (whose term order is not symmetric in z) always marches
    // upward; the first-half residuals (FR) take their two z terms in
    // stored-plane order, and the prolongation's corrections (PRO) are
    // evaluated at stored planes from a ring of stored coarse planes.
    auto pl = [&](int u) -> int { return dn ? g.nz - 1 - u : u; };
    auto plz = [&](int u) -> int { return dn ? max(g.nz - 2 - u, 0) : u; };
    auto at = [](const T* base, unsigned boff) -> T {
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
    };
    constexpr unsigned W = sizeof(T);

    // black phi_in of plane z, element i of a ring slot.  Loads are issued
    // unconditionally (an out-of-range element reads element 0 and is
    // zeroed on commit): branch-free issue lets the in-order load counter
    // leave them in flight across the step.
    auto src = [&](int z, int i, bool& ok) -> unsigned {
        const int rr_ = i / nh, kk = i - rr_ * nh;
        const int yy = y0 - 2 + rr_;
        const int x = 2 * kk + ((yy + z) & 1);
        ok = i < PR * nh && z >= 0 && z < g.nz && yy >= 0 && yy < g.ny && x < g.nx;
        return ok ? (unsigned)pl(z) * szb + (unsigned)yy * syb + (unsigned)x * W : 0u;
    };
    auto load_plane = [&](int z, auto corr_fn) {
        T* dst = P + (z % 3) * PR * nh;
        for (int i = tid; i < PR * nh; i += NT) {
            bool ok;
            const unsigned o = src(z, i, ok);
            dst[i] = corr_fn(z, i, ok ? at(pin, o) : (T)0);
        }
    };
    T pre[PMAX];
    unsigned pre_ok = 0;
    auto issue = [&](int z) {
        pre_ok = 0;
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            bool ok;
            const unsigned o = src(z, tid + j * NT, ok);
            pre[j] = at(pin, o);
            pre_ok |= ok ? (1u << j) : 0u;
        }
    };
    // PRO: the prolongation phi += P e (solver_3d.py:654-677) is applied to
    // phi_in's black values as they enter the ring (with nu2 > 0 the red
    // points' corrections are never read).  e comes from a ring of three
    // coarse planes, indexed by stored plane; stored fine plane f needs coarse
    // planes (f+1)>>1 .. (f+2)>>1, so the window moves up the coarse level
    // in an upward march and down it in a downward one.
    constexpr int cnx = NHMAX + 1;              // coarse ring row width (>= gc.nx)
    const int cr0 = (y0 - 2) >> 1;              // first coarse row of the ring
    auto cval = [&](int I, int J, int K) -> T { return Cr[((I % 3) * CRR + (J - cr0)) * cnx + K]; };
    // correction of ring element i of fine plane z (false: none)
    auto corr_of = [&](int z, int i, T& corr) -> bool {
        const int rr_ = i / nh, kk = i - rr_ * nh;
        const int yy = y0 - 2 + rr_;
        const int x = 2 * kk + ((yy + z) & 1);
        // a black point: x + yy + z is even (and x + yy + pl(z): mz + 2 is even)
        if (x < 2 || x > g.mx || yy < 2 || yy > g.my || z < 2 || z > g.mz) return false;
        corr = prolong3_corr_even<T>(cval, x, yy, pl(z));
        return true;
    };
    auto corrected = [&](int z, int i, T v) -> T {
        if constexpr (PRO) {
            T corr;
            if (corr_of(z, i, corr)) v = v + corr;
        }
        return v;
    };
    // The corrections of this thread's ring elements of plane z are computed
    // at the end of the step before its commit and parked in the element
    // itself (that ring slot is free by then); the commit adds the loaded
    // value.  Nothing stays live in registers across the step boundary.
    unsigned cpr_ok = 0;
    auto prep_corr = [&](int z) {
        if constexpr (PRO) {
            T* dst = P + (z % 3) * PR * nh;
            cpr_ok = 0;
            auto one = [&](int j) {
                const int i = tid + j * NT;
                T c;
                if (i < PR * nh && corr_of(z, i, c)) {
                    dst[i] = c;
                    cpr_ok |= 1u << j;
                }
            };
            if constexpr (sizeof(T) == 8 && WR) {
                // FP64 with rows of whole waves (two elements per thread): the
                // elements' coarse reads overlap (224 -> 220 us at 256^3, no
                // spills; the per-thread-pick variant would spill)
#pragma unroll
                for (int j = 0; j < PMAX; ++j) one(j);
            } else {
                // FP32: one element's coarse reads live at a time (unrolled,
                // the two-pair variants spill; unrolled by two, no faster)
#pragma unroll 1
                for (int j = 0; j < PMAX; ++j) one(j);
            }
        }
    };
    const T* pcb = PRO ? pc : pin;              // a valid base for the unconditional loads
    auto csrc = [&](int I, int i, bool& ok) -> unsigned {
        const int rr_ = i / cnx, kk = i - rr_ * cnx;
        const int J = cr0 + rr_;
        ok = i < CRR * cnx && kk < gc.nx && J < gc.ny && I >= 0 && I < gc.nz;
        return ok ? (unsigned)(((long)I * gc.sz + (long)J * gc.sy + kk) * (long)sizeof(T)) : 0u;
    };
    auto cload_plane = [&](int I) {
        T* dst = Cr + (I % 3) * CRR * cnx;
        for (int i = tid; i < CRR * cnx; i += NT) {
            bool ok;
            const unsigned o = csrc(I, i, ok);
            dst[i] = ok ? at(pcb, o) : (T)0;
        }
    };
    T cpre[CMAX];
    unsigned cpre_ok = 0;
    auto cissue = [&](int I) {
        cpre_ok = 0;
#pragma unroll
        for (int j = 0; j < CMAX; ++j) {
            bool ok;
            const unsigned o = csrc(I, tid + j * NT, ok);
            cpre[j] = at(pcb, o);
            cpre_ok |= ok ? (1u << j) : 0u;
        }
    };
    auto ccommit = [&](int I) {
        T* dst = Cr + (I % 3) * CRR * cnx;
#pragma unroll
        for (int j = 0; j < CMAX; ++j) {
            const int i = tid + j * NT;
            if (i < CRR * cnx) dst[i] = ((cpre_ok >> j) & 1u) ? cpre[j] : (T)0;
        }
    };
    auto cmax = [](int f) { return (f + 2) >> 1; };   // highest coarse plane of stored fine plane f
    auto cmin = [](int f) { return (f + 1) >> 1; };   // lowest (>= 0 for f >= -1)

    auto commit = [&](int z) {
        T* dst = P + (z % 3) * PR * nh;
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            const int i = tid + j * NT;
            if (i < PR * nh) {
                T v = ((pre_ok >> j) & 1u) ? pre[j] : (T)0;
                if constexpr (PRO) {
                    if ((cpr_ok >> j) & 1u) v = v + dst[i];
                }
                dst[i] = v;
            }
        }
    };
    auto Pv = [&](int z, int yy, int x) -> T {
        return P[((z % 3) * PR + (yy - y0 + 2)) * nh + (x >> 1)];
    };
    auto Rv = [&](int z, int yy, int x) -> T {
        return R[((z % 3) * RR + (yy - y0 + 1)) * nh + (x >> 1)];
    };
    using V = PairVec<T, 2 * NP>;
    auto atv = [](const T* base, unsigned boff) -> V {
        return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + boff);
    };
    auto load = [&](int z, PairGrp<T, NP>& c) {
        const unsigned q = (unsigned)pl(z) * szb + rowb;
        c.cxm = at(cx, q - W);
        c.cx = atv(cx, q);
        c.cy = atv(cy, q);
        c.cym = atv(cy, q - syb);
        c.cz = atv(cz, (unsigned)plz(z) * szb + rowb);
        c.rh = atv(rhs, q);
        c.dg = atv(cc, q);
    };
    // RES: the residual rhs - L phi_in at every unknown of the tile, in the
    // term order of resid3, summed as r^2 in FP64 (the first sweep of a cycle
    // evaluates the stop rule's residual of the cycle's input on the way)
    double rsum = 0.0;
    auto Rov = [&](int z, int yy, int x) -> T {
        return Ro[((z % 3) * RR + (yy - y0 + 1)) * nh + (x >> 1)];
    };
    // phi_in at the black point of the previous plane whose update runs one
    // step later, held by its thread like its coefficients (the ring slot of
    // that plane is refilled by the next step's commit before any barrier)
    auto Pbv = [&](int pp) -> T& { return Pb[pp * NT + tid]; };

    // PRO: the ring's newest coarse plane (upward: its highest, downward: its
    // lowest); the march planes z0 - 2 .. z0 + 1 need three
    int chave = 0;
    if constexpr (PRO) {
        if (dn) {
            chave = cmin(pl(z0 + 1));
            for (int I = chave; I <= cmax(pl(z0 - 2)); ++I) cload_plane(I);
        } else {
            chave = cmax(z0 + 1);
            for (int I = (z0 - 1) >> 1; I <= chave; ++I) cload_plane(I);
        }
        __syncthreads();
    }
    load_plane(z0 - 2, corrected);
    load_plane(z0 - 1, corrected);
    issue(z0);
    prep_corr(z0);
    PairGrp<T, NP> A, B;
    load(z0 - 1, A);
    V czp = atv(cz, (unsigned)plz(z0 - 2) * szb + rowb);   // cz of the plane below the current one
    // the black halves of the previous plane (their updates run one step
    // later) wait in LDS, private to the thread: no barrier needed
    auto H = [&](int pp, int j) -> T& { return Hb[(pp * 8 + j) * NT + tid]; };
    T red_prev[NP];
#pragma unroll
    for (int pp = 0; pp < NP; ++pp) red_prev[pp] = (T)0;
    // FR: the residual rhs - L phi at the red points (x + y + z odd) of the
    // swept level, in resid3's term order, written to rfo at res3_at.  The
    // residual of plane z needs the black values of planes z - 1, z and
    // z + 1, so it is formed over three steps: cc * phi and the z - 1
    // product when the red value is computed (step z), the four in-plane
    // terms and the z - 1 term after the black half of plane z (step z + 1),
    // the z + 1 term and rhs after the black half of plane z + 1 (step z + 2).  Only
    // the tile's rows y0 + 1 .. y0 + TY - 2 and planes z0 + 1 .. z1 - 2 are
    // done here: the other rows' and planes' neighbours belong to other
    // tiles (k_resid_edges3 covers those points).  Products are formed
    // early, subtractions in resid3's order, so the bits are resid3's.  On a
    // downward march the stored plane below is the march's next one: the z
    // term taken at step z + 2 comes first, and the product with the march's
    // previous plane waits one more step.
    //
    // Where a row of pairs lies within one wave (xs), the x terms come from
    // this thread's black values and the neighbouring lanes' at the end of
    // step z + 1, and the y terms after the first barrier of step z + 2 (the
    // y products of plane z wait in Bu, Bd until that step's black half), so
    // no third barrier is needed.  Otherwise (a row over several waves) the
    // four in-plane terms are taken after a third barrier of step z + 1, the
    // x neighbours from Bk.
    T t0h[NP];                                  // cc * phi of plane zs - 1's red points
    T carry[NP];                                // xs: plane zs - 2's x terms done; else plane zs - 1's z - 1 product
    T cxe = (T)0;                               // cx of plane zs - 1's red point at the thread's edge
#pragma unroll
    for (int pp = 0; pp < NP; ++pp) {
        t0h[pp] = (T)0;
        carry[pp] = (T)0;
    }
    const bool xs = gpr <= 64;                  // rows of pairs never straddle a wave
    // 0, 1: red rhs by plane parity; xs: 2, 3: the z - 1 product by plane
    // parity; otherwise 2: plane zs - 2's partial residual, 3 (downward
    // march): plane zs - 2's z - 1 product
    auto HF = [&](int j, int pp) -> T& { return Hf[(j * NP + pp) * NT + tid]; };
    const bool fr_row = FR && inner && r >= 2 && r <= TY - 1;
    auto fr_plane = [&](int u) { return u >= z0 + 1 && u <= z1 - 2; };

    // ec: which element of a pair is red on this step's plane, as a compile-time
    // constant (see the loop below)
    auto step = [&](int zs_in, PairGrp<T, NP>& cur, PairGrp<T, NP>& nxt, auto ec) {
        // the plane index is the same in every lane; from the first lane the
        // compiler keeps it and the ring slots derived from it in scalar registers
        const int zs = __builtin_amdgcn_readfirstlane(zs_in);
        // unconditional issue: planes up to z1 + 2 <= nz are requested (the
        // ring's guard zeroes plane nz; coefficient plane zs + 1 <= nz - 1)
        commit(zs + 1);
        if constexpr (PRO) {
            // coarse plane for the next step's commit (issued one step ago);
            // it replaces a plane no commit of this step reads
            if (dn) {
                const int want = cmin(pl(zs + 2));   // >= 0: pl(zs + 2) >= mz - z1 >= -1
                if (want < chave) {
                    ccommit(want);
                    chave = want;
                }
                cissue(cmin(pl(zs + 3)));   // a plane below 0 is not loaded, nor committed
            } else {
                const int want = cmax(zs + 2);
                if (want > chave) {
                    ccommit(want);
                    chave = want;
                }
                cissue(cmax(zs + 3));
            }
        }
        issue(zs + 2);
        load(zs + 1, nxt);
        // RES: phi_in of the thread's pairs on this plane (the ring fetched
        // these cache lines two steps ago); the barrier covers the latency
        V ph;
        if constexpr (RES) ph = atv(pin, (unsigned)pl(zs) * szb + rowb);
        __syncthreads();
        // ---- red on plane zs (x + y + z odd); pair pp covers x = 2(k+pp), 2(k+pp)+1
        const int e = decltype(ec)::at(y, zs);  // red element within a pair: 0 when y + zs is odd
        // FR, xs: plane zs - 2's red points (this plane's red element) take
        // their y terms, and on an upward march the z - 1 term
        T ptl[NP];
#pragma unroll
        for (int pp = 0; pp < NP; ++pp) ptl[pp] = (T)0;
        if constexpr (FR) {
            if (xs && fr_row && fr_plane(zs - 2)) {
#pragma unroll
                for (int pp = 0; pp < NP; ++pp) {
                    const int x = 2 * (k + pp) + e;
                    if (x <= g.mx) {
                        const int row = (r - 2) * nh + (x >> 1);
                        T rr = carry[pp] - Bu[row];
                        rr = rr - Bd[row];
                        if (!dn) rr = rr - HF(2 + (zs & 1), pp);
                        ptl[pp] = rr;
                    }
                }
            }
        }
        // element of pair pp: e ? odd : even (selects between constant indices;
        // a runtime index would move the vectors to scratch memory)
#define JSV_EL(vec, pp, odd) ((odd) ? (vec).v[2 * (pp) + 1] : (vec).v[2 * (pp)])
#define JSV_XM(c, pp, odd) ((odd) ? (c).cx.v[2 * (pp)] : ((pp) == 0 ? (c).cxm : (c).cx.v[2 * (pp) - 1]))
        T red[NP];
#pragma unroll
        for (int pp = 0; pp < NP; ++pp) {
            const int xr = 2 * (k + pp) + e;
            red[pp] = (T)0;
            if (unk && zs >= 2 && zs <= g.mz && xr <= g.mx) {
                const T a_zp = JSV_EL(cur.cz, pp, e);
                const T a_zm = JSV_EL(czp, pp, e);
                const T a_yp = JSV_EL(cur.cy, pp, e);
                const T a_ym = JSV_EL(cur.cym, pp, e);
                const T a_xp = JSV_EL(cur.cx, pp, e);
                const T a_xm = JSV_XM(cur, pp, e);
                // all ring reads first, then the reference's sum order
                const T p_zp = Pv(zs + 1, y, xr), p_zm = Pv(zs - 1, y, xr);
                const T p_yp = Pv(zs, y + 1, xr), p_ym = Pv(zs, y - 1, xr);
                const T p_xp = Pv(zs, y, xr + 1), p_xm = Pv(zs, y, xr - 1);
                T acc = a_zp * p_zp + a_zm * p_zm;
                acc = acc + a_yp * p_yp;
                acc = acc + a_ym * p_ym;
                acc = acc + a_xp * p_xp;
                acc = acc + a_xm * p_xm;
                acc = acc - JSV_EL(cur.rh, pp, e);
                red[pp] = acc / JSV_EL(cur.dg, pp, e);
                if constexpr (RES) {
                    if (inner && zs >= z0 && zs < z1) {
                        T rr = JSV_EL(cur.dg, pp, e) * JSV_EL(ph, pp, e);
                        rr = rr - a_xm * p_xm;
                        rr = rr - a_xp * p_xp;
                        rr = rr - a_ym * p_ym;
                        rr = rr - a_yp * p_yp;
                        rr = rr - a_zm * p_zm;
                        rr = rr - a_zp * p_zp;
                        rr = rr + JSV_EL(cur.rh, pp, e);
                        rsum += (double)rr * (double)rr;
                    }
                }
            }
            if (has) R[((zs % 3) * RR + r) * nh + k + pp] = red[pp];
            // RES: phi_in's red values of plane zs, for the black points' residuals
            if constexpr (RES)
                if (has) Ro[((zs % 3) * RR + r) * nh + k + pp] = ld ? JSV_EL(ph, pp, e) : (T)0;
        }
        __syncthreads();
        // ---- black on plane zb = zs - 1 at x = xr; write both colours of the pairs
        const int zb = zs - 1;
        T pmn[NP], bk[NP];                      // FR
#pragma unroll
        for (int pp = 0; pp < NP; ++pp) {
            pmn[pp] = (T)0;
            bk[pp] = (T)0;
        }
        if (inner && zb >= z0 && zb < z1) {
            T* orow = pout + ((long)pl(zb) * sz + (long)y * sy);
            // both colours of the thread's pairs leave as one vector store; a
            // pair reaching past the last unknown (x = mx + 1, a boundary
            // node) writes the zero that node holds
            PairVec<T, 2 * NP> ov;
#pragma unroll
            for (int pp = 0; pp < NP; ++pp) {
                const int xr = 2 * (k + pp) + e;
                T blk = (T)0;
                if (xr <= g.mx) {
                    const T r_zp = Rv(zb + 1, y, xr), r_zm = Rv(zb - 1, y, xr);
                    const T r_yp = Rv(zb, y + 1, xr), r_ym = Rv(zb, y - 1, xr);
                    const T r_xp = Rv(zb, y, xr + 1), r_xm = Rv(zb, y, xr - 1);
                    T acc = H(pp, 0) * r_zp + H(pp, 1) * r_zm;
                    acc = acc + H(pp, 2) * r_yp;
                    acc = acc + H(pp, 3) * r_ym;
                    acc = acc + H(pp, 4) * r_xp;
                    acc = acc + H(pp, 5) * r_xm;
                    acc = acc - H(pp, 6);
                    blk = acc / H(pp, 7);
                    if constexpr (RES) {
                        const T o_zp = Rov(zb + 1, y, xr), o_zm = Rov(zb - 1, y, xr);
                        const T o_yp = Rov(zb, y + 1, xr), o_ym = Rov(zb, y - 1, xr);
                        const T o_xp = Rov(zb, y, xr + 1), o_xm = Rov(zb, y, xr - 1);
                        T rr = H(pp, 7) * Pbv(pp);
                        rr = rr - H(pp, 5) * o_xm;
                        rr = rr - H(pp, 4) * o_xp;
                        rr = rr - H(pp, 3) * o_ym;
                        rr = rr - H(pp, 2) * o_yp;
                        rr = rr - H(pp, 1) * o_zm;
                        rr = rr - H(pp, 0) * o_zp;
                        rr = rr + H(pp, 6);
                        rsum += (double)rr * (double)rr;
                    }
                }
                if constexpr (FR) {
                    // this black row is y - y0 = r - 1: row r - 2 of Bk, the row
                    // above's Bu, the row below's Bd
                    const int col = k + pp;
                    if (!xs && r >= 2 && r <= TY - 1) Bk[(r - 2) * nh + col] = blk;
                    if (r <= TY - 2) Bu[(r - 1) * nh + col] = H(pp, 2) * blk;
                    if (r >= 3) Bd[(r - 3) * nh + col] = H(pp, 3) * blk;
                    bk[pp] = blk;
                    // the z - 1 product of plane zs's red point (same x)
                    const T pm = JSV_EL(czp, pp, e) * blk;
                    // plane zs - 2's red point (same x): its z + 1 term (cz of that
                    // point is held as this black point's z - 1 face) and rhs
                    if (fr_row && fr_plane(zs - 2) && xr <= g.mx) {
                        T rr = (xs ? ptl[pp] : HF(2, pp)) - H(pp, 1) * blk;
                        if (dn) rr = rr - HF(xs ? 2 + (zs & 1) : 3, pp);
                        rr = rr + HF(zs & 1, pp);
                        rfo[res3_at(g, pl(zs - 2), y, xr)] = rr;
                    }
                    // (xs: the slot of plane zs - 2's product, read just above)
                    if (xs) HF(2 + (zs & 1), pp) = pm;
                    else pmn[pp] = pm;
                }
                const T rp = ((xr ^ 1) <= g.mx) ? red_prev[pp] : (T)0;
                ov.v[2 * pp] = e ? rp : blk;
                ov.v[2 * pp + 1] = e ? blk : rp;
            }
            *reinterpret_cast<PairVec<T, 2 * NP>*>(orow + 2 * k) = ov;
        } else if (FR && has && y == g.my + 1 && r >= 3 && r <= TY && zb >= z0 && zb < z1) {
            // the boundary row above the last unknown row: the y + 1 terms of
            // that row's red points multiply cy by this row's zero phi
#pragma unroll
            for (int pp = 0; pp < NP; ++pp) Bd[(r - 3) * nh + k + pp] = H(pp, 3) * (T)0;
        }
        if constexpr (FR) {
            // plane zs - 1's red points at x = 2(k + pp) + 1 - e.  Their cx faces
            // are held by this thread's black halves of the same plane (cxe at
            // its edge).
            auto faces = [&](int pp, T& fxm, T& fxp) {
                if (e == 1) {
                    fxp = H(pp, 5);
                    fxm = pp > 0 ? H(pp - 1, 4) : cxe;
                } else {
                    fxm = H(pp, 4);
                    fxp = pp + 1 < NP ? H(pp + 1, 5) : cxe;
                }
            };
            if (xs) {
                // cc * phi and the x terms: the x neighbours are this row's black
                // values of the plane, this thread's or the lanes' beside it (x =
                // 1 and x = mx + 1 are boundary nodes)
                const T bl = __shfl_up(bk[NP - 1], 1, 64), br = __shfl_down(bk[0], 1, 64);
                if (fr_row && fr_plane(zs - 1)) {
#pragma unroll
                    for (int pp = 0; pp < NP; ++pp) {
                        const int x = 2 * (k + pp) + 1 - e;
                        if (x <= g.mx) {
                            T fxm, fxp, bxm, bxp;
                            faces(pp, fxm, fxp);
                            if (e == 1) {               // x + 1 is this pair's black point
                                bxp = bk[pp];
                                bxm = pp > 0 ? bk[pp - 1] : (k == 1 ? (T)0 : bl);
                            } else {                    // x - 1 is this pair's black point
                                bxm = bk[pp];
                                bxp = pp + 1 < NP ? bk[pp + 1] : br;
                            }
                            T rr = t0h[pp] - fxm * bxm;
                            rr = rr - fxp * bxp;
                            carry[pp] = rr;
                        }
                    }
                }
            } else {
                __syncthreads();                // every row's black values of plane zs - 1 are in Bk, Bu, Bd
                // cc * phi, the in-plane terms, the z - 1 term
                if (fr_row && fr_plane(zs - 1)) {
#pragma unroll
                    for (int pp = 0; pp < NP; ++pp) {
                        const int x = 2 * (k + pp) + 1 - e;
                        if (x <= g.mx) {
                            T fxm, fxp;
                            faces(pp, fxm, fxp);
                            const int row = (r - 2) * nh;
                            const T bxm = Bk[row + ((x - 1) >> 1)], bxp = Bk[row + ((x + 1) >> 1)];
                            const T pym = Bu[row + (x >> 1)], pyp = Bd[row + (x >> 1)];
                            T rr = t0h[pp] - fxm * bxm;
                            rr = rr - fxp * bxp;
                            rr = rr - pym;
                            rr = rr - pyp;
                            if (dn) HF(3, pp) = carry[pp];
                            else rr = rr - carry[pp];
                            HF(2, pp) = rr;
                        }
                    }
                }
            }
        }
        // ---- keep plane zs's black halves (the points that are not red)
#pragma unroll
        for (int pp = 0; pp < NP; ++pp) {
            const int b1 = 1 - e;               // black element within the pair
            H(pp, 0) = JSV_EL(cur.cz, pp, b1);
            H(pp, 1) = JSV_EL(czp, pp, b1);
            H(pp, 2) = JSV_EL(cur.cy, pp, b1);
            H(pp, 3) = JSV_EL(cur.cym, pp, b1);
            H(pp, 4) = JSV_EL(cur.cx, pp, b1);
            H(pp, 5) = JSV_XM(cur, pp, b1);
            H(pp, 6) = JSV_EL(cur.rh, pp, b1);
            H(pp, 7) = JSV_EL(cur.dg, pp, b1);
            if constexpr (RES) Pbv(pp) = JSV_EL(ph, pp, b1);
            if constexpr (FR) {
                HF(zs & 1, pp) = JSV_EL(cur.rh, pp, e);
                t0h[pp] = JSV_EL(cur.dg, pp, e) * red[pp];
                if (!xs) carry[pp] = pmn[pp];
            }
            red_prev[pp] = red[pp];
        }
        if constexpr (FR) cxe = e ? cur.cx.v[2 * NP - 1] : cur.cxm;
#undef JSV_EL
#undef JSV_XM
        czp = cur.cz;
        // coarse planes up to cmax(zs + 2) are in the ring since this step's
        // first barrier; ring slot (zs + 2) % 3 last held plane zs - 1, read
        // by this step's red half before the second barrier
        prep_corr(zs + 2);
    };
    // The red element of a pair is 0 where y + z is odd, 1 where it is even, so
    // it alternates with the step.  When a row of pairs is a whole number of
    // waves (gpr a multiple of 64: every wave holds one row, or idle threads
    // only), the branch on the row's parity is uniform within each wave and
    // each path has the element picks resolved at compile time; both paths run
    // the same barriers in the same order (s_barrier counts the waves of the
    // workgroup).  Otherwise (WR false) a wave holds several rows and the
    // element is computed per thread.  The host sets WR from gpr.
    if constexpr (WR) {
        // the branch sits inside the loop, so the loop counter stays uniform
        const bool odd_first = ((y + z0 - 1) & 1) != 0;
        for (int zs = z0 - 1; zs <= z1; zs += 2) {
            if (odd_first) step(zs, A, B, PairElem<0>{});
            else step(zs, A, B, PairElem<1>{});
            if (zs + 1 <= z1) {
                if (odd_first) step(zs + 1, B, A, PairElem<1>{});
                else step(zs + 1, B, A, PairElem<0>{});
            }
        }
    } else {
        for (int zs = z0 - 1; zs <= z1; zs += 2) {
            step(zs, A, B, PairElemRT{});
            if (zs + 1 <= z1) step(zs + 1, B, A, PairElemRT{});
        }
    }
    if constexpr (RES) {
        // one partial sum per workgroup, in a fixed order
        const double v = wave_sum(rsum);
        __syncthreads();                        // the rings are no longer read
        double* sh = reinterpret_cast<double*>(jsv_dyn_smem);
        if ((tid & 63) == 0) sh[tid >> 6] = v;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int i = 0; i < NT / 64; ++i) t += sh[i];
            res_part[blockIdx.x] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// Residual rhs - L phi at one point (solver_2d.py:637, solver_3d.py:642), in the
// reference's term order.
// ---------------------------------------------------------------------------
template <typename T, typename TP>
__device__ __forceinline__ T resid2(const TP* __restrict__ phi, const TP* __restrict__ rhs,
                                    const T* __restrict__ cx, const T* __restrict__ cy,
                                    const T* __restrict__ cc, const T* __restrict__ drr,
                                    const T* __restrict__ dlr, long p, long sy)
{
    T r = cc[p] * (T)phi[p];
    r = r - cx[p - 1] * (T)phi[p - 1];
    r = r - cx[p] * (T)phi[p + 1];
    r = r - cy[p - sy] * (T)phi[p - sy];
    r = r - cy[p] * (T)phi[p + sy];
    r = r + (T)rhs[p];
    if (drr != nullptr) {
        T off = drr[p - sy - 1] * (T)phi[p - sy - 1] + dlr[p - sy + 1] * (T)phi[p - sy + 1];
        off = off + dlr[p] * (T)phi[p + sy - 1];
        off = off + drr[p] * (T)phi[p + sy + 1];
        r = r - off;
    }
    return r;
}

template <typename T, typename TP>
__device__ __forceinline__ T resid3(const TP* __restrict__ phi, const TP* __restrict__ rhs,
                                    const T* __restrict__ cx, const T* __restrict__ cy,
                                    const T* __restrict__ cz, const T* __restrict__ cc,
                                    long p, long sy, long sz)
{
    T r = cc[p] * (T)phi[p];
    r = r - cx[p - 1] * (T)phi[p - 1];
    r = r - cx[p] * (T)phi[p + 1];
    r = r - cy[p - sy] * (T)phi[p - sy];
    r = r - cy[p] * (T)phi[p + sy];
    r = r - cz[p - sz] * (T)phi[p - sz];
    r = r - cz[p] * (T)phi[p + sz];
    r = r + (T)rhs[p];
    return r;
}

// ---------------------------------------------------------------------------
// Defect + partial full-weighting restriction (solver_2d.py:632-647,
// solver_3d.py:638-652): one thread per coarse unknown; the residuals it needs
// (2D: centre + 4 corners, 3D: centre + 12 edge points, all on the first RBGS
// half) are evaluated in registers, never stored.  Coarse boundary nodes and
// ghosts are not written (zero from allocation).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T restrict2_at(const T* __restrict__ phi, const T* __restrict__ rhs,
                                          const T* __restrict__ cx, const T* __restrict__ cy,
                                          const T* __restrict__ cc, const T* __restrict__ drr,
                                          const T* __restrict__ dlr, long c, long sy)
{
    // c: the fine point 2I-1 under coarse unknown I (problem offset included)
    T s = resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, c + sy + 1, sy)
        + resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, c + sy - 1, sy);
    s = s + resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, c - sy + 1, sy);
    s = s + resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, c - sy - 1, sy);
    const T rcen = resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, c, sy);
    return (T)0.0625 * s + (T)0.25 * rcen;
}

template <typename T>
__device__ __forceinline__ T restrict3_at(const T* __restrict__ phi, const T* __restrict__ rhs,
                                          const T* __restrict__ cx, const T* __restrict__ cy,
                                          const T* __restrict__ cz, const T* __restrict__ cc, long c,
                                          long sy, long sz)
{
#define JSV_R(o) resid3<T, T>(phi, rhs, cx, cy, cz, cc, c + (o), sy, sz)
    T e = JSV_R(sy + 1) + JSV_R(sy - 1);
    e = e + JSV_R(-sy + 1);
    e = e + JSV_R(-sy - 1);
    e = e + JSV_R(sz + 1);
    e = e + JSV_R(sz - 1);
    e = e + JSV_R(-sz + 1);
    e = e + JSV_R(-sz - 1);
    e = e + JSV_R(sz + sy);
    e = e + JSV_R(sz - sy);
    e = e + JSV_R(-sz + sy);
    e = e + JSV_R(-sz - sy);
    const T rcen = JSV_R(0);
#undef JSV_R
    return (T)0.03125 * e + (T)0.125 * rcen;
}

// The first RBGS half of the level a restriction produces, run inside the
// restriction kernel.  That level's phi is zero when its first visit starts,
// so the first half's new value at a point needs only the point's own rhs
// (just computed by the same thread) and coefficients: the expression of
// rb2_point / rb3_point with every neighbour value an exact zero (products
// kept, so NaN and Inf coefficients propagate as in the separate launch).
// cc == nullptr: disabled.
template <typename T>
struct RedZ {
    const T *cx, *cy, *cz, *cc;
};

template <typename T>
__device__ __forceinline__ T rb2_from_zero(T rc, const RedZ<T>& z, long p, long sy)
{
    const T o = (T)0;
    T acc = z.cy[p] * o + z.cy[p - sy] * o;
    acc = acc + z.cx[p] * o;
    acc = acc + z.cx[p - 1] * o;
    acc = acc - rc;
    return acc / z.cc[p];
}

template <typename T>
__device__ __forceinline__ T rb3_from_zero(T rc, const RedZ<T>& z, long p, long sy, long sz)
{
    const T o = (T)0;
    T acc = z.cz[p] * o + z.cz[p - sz] * o;
    acc = acc + z.cy[p] * o;
    acc = acc + z.cy[p - sy] * o;
    acc = acc + z.cx[p] * o;
    acc = acc + z.cx[p - 1] * o;
    acc = acc - rc;
    return acc / z.cc[p];
}

// Zero the coarse phi that the cycle visits next, in the restriction kernel
// instead of a separate fill: the thread of coarse unknown (I, J, K) clears it
// and the rim points whose nearest unknown it is, so the whole array ends zero.
template <typename T>
__device__ __forceinline__ void zero_with_rim(T* __restrict__ a, const Geo& gc, long base, int dim, int I,
                                              int J, int K)
{
    const int z_lo = (dim == 3 && I == 2) ? 0 : I, z_hi = (dim == 3 && I == gc.mz) ? gc.mz + 2 : I;
    const int y_lo = J == 2 ? 0 : J, y_hi = J == gc.my ? gc.my + 2 : J;
    const int x_lo = K == 2 ? 0 : K, x_hi = K == gc.mx ? gc.mx + 2 : K;
    for (int z = z_lo; z <= z_hi; ++z)
        for (int y = y_lo; y <= y_hi; ++y)
            for (int x = x_lo; x <= x_hi; ++x)
                a[base + (dim == 3 ? (long)z * gc.sz : 0) + (long)y * gc.sy + x] = (T)0;
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_restrict2(T* __restrict__ rc, const T* __restrict__ phi, const T* __restrict__ rhs,
            const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cc,
            const T* __restrict__ drr, const T* __restrict__ dlr, Geo gf, Geo gc,
            const int* __restrict__ act, T* __restrict__ zphi, RedZ<T> rz)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int J = 2 + blockIdx.y * BY + threadIdx.y;
    const int K = 2 + blockIdx.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    const long c = (long)b * gf.sz + (long)(2 * J - 1) * gf.sy + (2 * K - 1);
    const long q = (long)b * gc.sz + (long)J * gc.sy + K;
    const T v = restrict2_at(phi, rhs, cx, cy, cc, drr, dlr, c, gf.sy);
    rc[q] = v;
    if (zphi != nullptr) zero_with_rim(zphi, gc, (long)b * gc.sz, 2, 0, J, K);
    if (rz.cc != nullptr && ((J + K) & 1) == 0) zphi[q] = rb2_from_zero(v, rz, q, gc.sy);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_restrict3(T* __restrict__ rc, const T* __restrict__ phi, const T* __restrict__ rhs,
            const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
            const T* __restrict__ cc, Geo gf, Geo gc, T* __restrict__ zphi, RedZ<T> rz,
            const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const int I = 2 + blockIdx.z;
    const int J = 2 + blockIdx.y * BY + threadIdx.y;
    const int K = 2 + blockIdx.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    const long c = (long)(2 * I - 1) * gf.sz + (long)(2 * J - 1) * gf.sy + (2 * K - 1);
    const long q = (long)I * gc.sz + (long)J * gc.sy + K;
    const T v = restrict3_at(phi, rhs, cx, cy, cz, cc, c, gf.sy, gf.sz);
    rc[q] = v;
    if (zphi != nullptr) zero_with_rim(zphi, gc, 0, 3, I, J, K);
    if (rz.cc != nullptr && ((I + J + K) & 1) == 1) zphi[q] = rb3_from_zero(v, rz, q, gc.sy, gc.sz);
}

// Two-pass restriction for large levels: pass 1 stores the residual at every
// point of the first RBGS half (exactly the points the partial full-weighting
// reads) once; pass 2 combines them per coarse unknown in the reference's
// order.  Same bits as the one-pass kernels above, without evaluating each
// edge residual four times.

// RB rows per block: tall blocks on large levels, so the rows above and below
// a block (read for the y-neighbours) are few against the rows it owns.
template <typename T, int RB = BY>
__global__ void __launch_bounds__(BX * RB)
k_resid_first3(T* __restrict__ r, const T* __restrict__ phi, const T* __restrict__ rhs,
               const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
               const T* __restrict__ cc, Geo g, int remap, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    // remap: XCD-slab block order (pays on levels the Infinity Cache holds;
    // on larger levels the launch order streams HBM better)
    const uint3 bk = remap ? xcd_block() : make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
    const int y = 2 + bk.y * RB + threadIdx.y;
    const int z = 2 + bk.z;
    if (y > g.my) return;
    const int x = 2 + ((1 + y + z) & 1) + 2 * (bk.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long p = (long)z * g.sz + (long)y * g.sy + x;
    r[res3_at(g, z, y, x)] = resid3<T, T>(phi, rhs, cx, cy, cz, cc, p, g.sy, g.sz);
}

// The first-half residuals a pre-smoothing pair sweep with FR leaves out: the
// first and last row (if it exists) of each of its ty-row tiles, on every
// plane, and the first and last plane of each z-chunk (the other rows).  The
// edge rows come in adjacent pairs at the tile boundaries (last row of tile
// j - 1, first row of tile j), the edge planes likewise: a workgroup takes one
// such pair on one plane (one row), its two halves a row (plane) each, and
// consecutive workgroups step the plane (row), so the rows and planes they
// share are read once (XCD k runs the k-th eighth of that order, so they
// share its L2).  Blocks [0, (ytiles + 1)(mz - 1)) take the rows;
// resid3 at each red point, as k_resid_first3.
// dn: the sweep marched downward (march plane u is stored plane mz + 2 - u).
constexpr int EDGE_THREADS = 256;
template <typename T>
__global__ void __launch_bounds__(EDGE_THREADS)
k_resid_edges3(T* __restrict__ r, const T* __restrict__ phi, const T* __restrict__ rhs,
               const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cz,
               const T* __restrict__ cc, Geo g, int ty, int zchunk, int ytiles, int dn,
               const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const int h = threadIdx.x >> 7, lane = threadIdx.x & 127;
    const int nA = (ytiles + 1) * (g.mz - 1);
    int b = xcd_tile((int)blockIdx.x, (int)gridDim.x), z, y;
    if (b < nA) {
        const int j = b / (g.mz - 1);
        z = 2 + (b - j * (g.mz - 1));
        if (h == 0) {
            y = 2 + j * ty - 1;                   // last row of tile j - 1
            if (j == 0 || y > g.my) return;
        } else {
            y = 2 + j * ty;                       // first row of tile j
            if (j == ytiles) return;
        }
    } else {
        b -= nA;
        const int j = b / (g.my - 1);
        y = 2 + (b - j * (g.my - 1));
        const int ry = (y - 2) % ty;
        if (ry == 0 || ry == ty - 1) return;      // a tile edge row: the first blocks have it
        const int zchunks = (g.mz - 1 + zchunk - 1) / zchunk;
        if (h == 0) {
            if (j == 0) return;
            z = min(2 + j * zchunk - 1, g.mz);    // last plane of chunk j - 1
            if (z == 2 + (j - 1) * zchunk) return;   // also its first: slot j - 1 has it
        } else {
            z = 2 + j * zchunk;                   // first plane of chunk j
            if (j == zchunks) return;
        }
    }
    if (dn) z = g.mz + 2 - z;
    const int x0 = 2 + ((1 + y + z) & 1);         // x + y + z odd
    for (int i = lane; x0 + 2 * i <= g.mx; i += 128) {
        const int x = x0 + 2 * i;
        const long p = (long)z * g.sz + (long)y * g.sy + x;
        r[res3_at(g, z, y, x)] = resid3<T, T>(phi, rhs, cx, cy, cz, cc, p, g.sy, g.sz);
    }
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_restrict_gather3(T* __restrict__ rc, const T* __restrict__ r, Geo gf, Geo gc, T* __restrict__ zphi,
                   RedZ<T> rz, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const uint3 bk = xcd_block();
    const int I = 2 + bk.z;
    const int J = 2 + bk.y * BY + threadIdx.y;
    const int K = 2 + bk.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    // residual at fine offset (dz, dy, dx) from the centre (2I-1, 2J-1, 2K-1):
    // res3_at with the row offsets formed once, in 32 bits (the compact array
    // holds fewer than 2^32 elements; the host checks), and (x-2)/2 = K-1 for
    // dx = +1, K-2 for dx = 0 and dx = -1
    const unsigned mxh = (unsigned)(gf.mx >> 1), pz = (unsigned)(gf.my - 1) * mxh;
    const unsigned c0 = ((unsigned)(2 * I - 3) * (unsigned)(gf.my - 1) + (unsigned)(2 * J - 3)) * mxh;
    auto R = [&](int dz, int dy, int dx) -> T {
        return r[c0 + (unsigned)dz * pz + (unsigned)dy * mxh + (unsigned)(dx > 0 ? K - 1 : K - 2)];
    };
    T e = R(0, 1, 1) + R(0, 1, -1);
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
    const long q = (long)I * gc.sz + (long)J * gc.sy + K;
    const T v = (T)0.03125 * e + (T)0.125 * R(0, 0, 0);
    rc[q] = v;
    if (zphi != nullptr) zero_with_rim(zphi, gc, 0, 3, I, J, K);
    if (rz.cc != nullptr && ((I + J + K) & 1) == 1) zphi[q] = rb3_from_zero(v, rz, q, gc.sy, gc.sz);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_resid_first2(T* __restrict__ r, const T* __restrict__ phi, const T* __restrict__ rhs,
               const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ cc,
               const T* __restrict__ drr, const T* __restrict__ dlr, Geo g, const int* __restrict__ act)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    if (y > g.my) return;
    const int x = 2 + (y & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (x > g.mx) return;
    const long p = (long)b * g.sz + (long)y * g.sy + x;
    r[p] = resid2<T, T>(phi, rhs, cx, cy, cc, drr, dlr, p, g.sy);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_restrict_gather2(T* __restrict__ rc, const T* __restrict__ r, Geo gf, Geo gc,
                   const int* __restrict__ act, T* __restrict__ zphi, RedZ<T> rz)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int J = 2 + blockIdx.y * BY + threadIdx.y;
    const int K = 2 + blockIdx.x * BX + threadIdx.x;
    if (J > gc.my || K > gc.mx) return;
    const long sy = gf.sy;
    const long c = (long)b * gf.sz + (long)(2 * J - 1) * sy + (2 * K - 1);
    T s = r[c + sy + 1] + r[c + sy - 1];
    s = s + r[c - sy + 1];
    s = s + r[c - sy - 1];
    const long q = (long)b * gc.sz + (long)J * gc.sy + K;
    const T v = (T)0.0625 * s + (T)0.25 * r[c];
    rc[q] = v;
    if (zphi != nullptr) zero_with_rim(zphi, gc, (long)b * gc.sz, 2, 0, J, K);
    if (rz.cc != nullptr && ((J + K) & 1) == 0) zphi[q] = rb2_from_zero(v, rz, q, gc.sy);
}

// ---------------------------------------------------------------------------
// Prolongation + add (solver_2d.py:612, 649-668; solver_3d.py:618, 654-677).
// Fine index i <-> coarse (i+1)/2 for odd i; an even i lies between coarse
// i/2 and i/2+1.  The interpolations the reference skips when nu2 > 0 (points
// of the first RBGS half, overwritten before they are read) are skipped alike.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_prolong2(T* __restrict__ phi, const T* __restrict__ pcb, Geo gf, Geo gc, int full,
           const int* __restrict__ act)
{
    const int b = blockIdx.z;
    if (act != nullptr && act[b] == 0) return;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gf.my || x > gf.mx) return;
    T corr;
    if (!prolong2_corr(pcb + (long)b * gc.sz, gc.sy, x, y, full, corr)) return;
    const long p = (long)b * gf.sz + (long)y * gf.sy + x;
    phi[p] = phi[p] + corr;
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_prolong3(T* __restrict__ phi, const T* __restrict__ pc, Geo gf, Geo gc, int full, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    const int z = 2 + blockIdx.z;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > gf.my || x > gf.mx) return;
    T corr;
    if (!prolong3_corr(pc, gc.sy, gc.sz, x, y, z, full, corr)) return;
    const long p = (long)z * gf.sz + (long)y * gf.sy + x;
    phi[p] = phi[p] + corr;
}

// ---------------------------------------------------------------------------
// Coarse-level program.  Below ~16^3 (3D) / 64^2 (2D) unknowns a level's
// smoothing half, restriction or prolongation is a few thousand points: as
// separate launches each costs a launch plus a fill/drain of the whole GPU for
// microseconds of work.  One workgroup per problem instead runs the whole
// coarse sub-cycle from a host-built op list, with a workgroup barrier between
// ops.  The per-point arithmetic is the same device functions the
// launch-per-op kernels use, so results are bit-identical.
//
// Latency, not bandwidth, bounds these ops, so the coarsest levels live in
// LDS for the whole launch: levels [lds_from, lds_to) are addressed with LDS
// instructions (each op body is instantiated per storage of the two levels it
// touches, so the compiler sees LDS pointers, not generic ones).  Their
// coefficients arrive in one pipelined copy from a prebuilt image that holds
// them in LDS order; their rhs and phi are produced inside the program (the
// first op zeroes the start level's phi; each restriction writes the next
// level's rhs and zeroes its phi), except the start level's rhs, copied in,
// and its phi, copied back at the end.  Runs of ops on LDS levels that touch
// no more points than one wave has lanes execute in wave 0 alone, in order,
// without workgroup barriers (LDS operations of one wave complete in order).
// ---------------------------------------------------------------------------
template <typename T>
struct LvPtr {
    T* phi;
    T* rhs;
    const T* cx;
    const T* cy;
    const T* cz;
    const T* cc;
    const T* drr;
    const T* dlr;
    Geo g;
    // LDS copy of this level, used when the level lies in the staged range of
    // the launch: element offsets of phi, rhs, cx, cy, cz, cc, drr, dlr (-1:
    // absent), each a compact array over indices 1..m+1 per axis (the ghost
    // layers 0 and m+2 are never read) of lsz elements per plane, lsy per row
    // and ln in total
    int lds[8];
    int lsy, lsz, ln;
};

enum CoarseOpKind : int {
    CO_RB = 0,        // one RBGS half: a = parity
    CO_RB_DIAG = 1,   // one D_xy sub-colour: a = lx, b = ly
    CO_RESTRICT = 2,  // defect of `level` restricted into rhs of level + 1; b = 1: also phi of level + 1 := 0;
                      // a = 1 (2D, with b = 1): level + 1's first RBGS half from that zero phi as well
    CO_ZERO = 3,      // phi of `level` := 0 (whole array incl. rim)
    CO_PROLONG = 4,   // phi of `level` += P phi of level + 1; a = full
    CO_RB_PAIR = 5,   // both RBGS halves, parity a first (one-wave runs only)
    CO_PRO_RB = 6,    // 2D: the prolongation (full = 0) inside the RBGS half of parity a; b = 1: then
                      // the other half (one-wave runs only)
};

struct CoarseOp {
    int kind, level, a, b;
    int run;          // > 0: first of `run` consecutive ops that wave 0 runs alone
    int reg;          // (first op of a run) the run keeps its levels' points in registers
};

// An op as the kernel reads it: one 32-bit word (kind 3 bits, level 5, a 1,
// b 1, run 21, reg 1), so the whole list sits in LDS next to the levels.
__host__ __device__ __forceinline__ unsigned coarse_pack(const CoarseOp& op)
{
    return (unsigned)op.kind | ((unsigned)op.level << 3) | ((unsigned)op.a << 8) | ((unsigned)op.b << 9) |
           ((unsigned)op.run << 10) | ((unsigned)op.reg << 31);
}
__device__ __forceinline__ CoarseOp coarse_unpack(unsigned w)
{
    return {(int)(w & 7u), (int)((w >> 3) & 31u), (int)((w >> 8) & 1u), (int)((w >> 9) & 1u),
            (int)((w >> 10) & 0x1fffffu), (int)(w >> 31)};
}
constexpr int COARSE_MAX_OPS = 8192;   // longer programs run as separate launches
constexpr int COARSE_THREADS = 1024;
constexpr int COARSE_WAVE = 64;

// Array pointers of one level and the map from padded coordinates to their
// element index, p = z * sz + y * sy + x - org: the LDS copy (IN_LDS, compact
// strides, org = the offset of point (1, 1, 1)) or the global arrays at
// problem offset bo (padded strides, org = 0).  zb / zn: the whole array, for
// zeroing.
template <typename T, bool IN_LDS>
struct LvView {
    T* phi;
    T* rhs;
    const T* cx;
    const T* cy;
    const T* cz;
    const T* cc;
    const T* drr;
    const T* dlr;
    bool od;
    long sy, sz, org;
    int zn;
    __device__ __forceinline__ LvView(const LvPtr<T>& L, T* sm, long bo, int dim)
    {
        if constexpr (IN_LDS) {
            phi = sm + L.lds[0];
            rhs = sm + L.lds[1];
            cx = sm + L.lds[2];
            cy = sm + L.lds[3];
            cz = sm + (L.lds[4] >= 0 ? L.lds[4] : 0);
            cc = sm + L.lds[5];
            od = L.lds[6] >= 0;
            drr = sm + (od ? L.lds[6] : 0);
            dlr = sm + (od ? L.lds[7] : 0);
            sy = L.lsy;
            sz = L.lsz;
            org = (dim == 3 ? sz : 0) + sy + 1;
            zn = L.ln;
        } else {
            phi = L.phi + bo;
            rhs = L.rhs + bo;
            cx = L.cx + bo;
            cy = L.cy + bo;
            cz = L.cz != nullptr ? L.cz + bo : L.cx + bo;
            cc = L.cc + bo;
            od = L.drr != nullptr;
            drr = od ? L.drr + bo : L.cx + bo;
            dlr = od ? L.dlr + bo : L.cx + bo;
            sy = L.g.sy;
            sz = L.g.sz;
            org = 0;
            zn = (int)(dim == 2 ? L.g.sz : L.g.n);
        }
    }
    __device__ __forceinline__ long at(int z, int y, int x) const { return (long)z * sz + (long)y * sy + x - org; }
    __device__ __forceinline__ long at(int y, int x) const { return (long)y * sy + x - org; }
};

// Point updates of the coarse program.  Its small ops leave one lane one
// point, so a point's latency is the op's: every operand is loaded before any
// arithmetic (one LDS round trip instead of several), and the arithmetic is
// rb2_point / rb3_point / restrict2_at's own, in the same order.
template <typename T>
__device__ __forceinline__ void rb2_point_pre(T* __restrict__ phi, const T* __restrict__ rhs,
                                              const T* __restrict__ cx, const T* __restrict__ cy,
                                              const T* __restrict__ cc, long p, long sy)
{
    const T a0 = cy[p], a1 = phi[p + sy], a2 = cy[p - sy], a3 = phi[p - sy], a4 = cx[p], a5 = phi[p + 1],
            a6 = cx[p - 1], a7 = phi[p - 1], r = rhs[p], q = cc[p];
    __builtin_amdgcn_sched_barrier(0);
    T acc = a0 * a1 + a2 * a3;
    acc = acc + a4 * a5;
    acc = acc + a6 * a7;
    acc = acc - r;
    phi[p] = acc / q;
}

template <typename T>
__device__ __forceinline__ void rb3_point_pre(T* __restrict__ phi, const T* __restrict__ rhs,
                                              const T* __restrict__ cx, const T* __restrict__ cy,
                                              const T* __restrict__ cz, const T* __restrict__ cc, long p,
                                              long sy, long sz)
{
    const T a0 = cz[p], a1 = phi[p + sz], a2 = cz[p - sz], a3 = phi[p - sz], a4 = cy[p], a5 = phi[p + sy],
            a6 = cy[p - sy], a7 = phi[p - sy], a8 = cx[p], a9 = phi[p + 1], a10 = cx[p - 1], a11 = phi[p - 1],
            r = rhs[p], q = cc[p];
    __builtin_amdgcn_sched_barrier(0);
    T acc = a0 * a1 + a2 * a3;
    acc = acc + a4 * a5;
    acc = acc + a6 * a7;
    acc = acc + a8 * a9;
    acc = acc + a10 * a11;
    acc = acc - r;
    phi[p] = acc / q;
}

// restrict2_at without D_xy terms: the operands of resid2 at the five points
template <typename T>
__device__ __forceinline__ T restrict2_at_pre(const T* __restrict__ phi, const T* __restrict__ rhs,
                                              const T* __restrict__ cx, const T* __restrict__ cy,
                                              const T* __restrict__ cc, long c, long sy)
{
    struct Ops {
        T v[11];
    };
    auto load = [&](long p) {
        return Ops{{cc[p], phi[p], cx[p - 1], phi[p - 1], cx[p], phi[p + 1], cy[p - sy], phi[p - sy], cy[p],
                    phi[p + sy], rhs[p]}};
    };
    auto res = [](const Ops& o) {
        T r = o.v[0] * o.v[1];
        r = r - o.v[2] * o.v[3];
        r = r - o.v[4] * o.v[5];
        r = r - o.v[6] * o.v[7];
        r = r - o.v[8] * o.v[9];
        return r + o.v[10];
    };
    const Ops a = load(c + sy + 1), b = load(c + sy - 1), d = load(c - sy + 1), e = load(c - sy - 1), f = load(c);
    __builtin_amdgcn_sched_barrier(0);
    T s = res(a) + res(b);
    s = s + res(d);
    s = s + res(e);
    const T rcen = res(f);
    return (T)0.0625 * s + (T)0.25 * rcen;
}

// Point i of an op's index space: every level enumerates power-of-two extents
// (resolutions are powers of two) with shifts; points past the last row /
// plane are skipped by the caller.
struct Ix2 {
    int hi, lo;
};
__device__ __forceinline__ Ix2 split2(int i, int nlo, int llo)
{
    return {i >> llo, i & (nlo - 1)};
}

// One op over points t0, t0 + nt, ...; FL / CL: level / level + 1 in LDS.
template <typename T, int dim, bool FL, bool CL>
__device__ __forceinline__ void coarse_op(const CoarseOp& op, const LvPtr<T>* __restrict__ lv, T* sm, int b,
                                          int t0, int nt)
{
    const LvPtr<T>& LF = lv[op.level];
    const Geo g = LF.g;
    const LvView<T, FL> F(LF, sm, (dim == 2 && !FL) ? (long)b * g.sz : 0, dim);
    // loop extents are the resolutions: padded past the last unknown to a power of two
    auto lg = [](int v) { return v > 0 ? __builtin_ctz((unsigned)v) : 0; };
    if (op.kind == CO_RB || op.kind == CO_RB_PAIR || op.kind == CO_PRO_RB) {
        const int xh = g.mx / 2, ny = g.my, lxh = lg(xh), lny = lg(ny);
        auto half = [&](int par) {
            if constexpr (dim == 3) {
                const int cnt = xh * ny * g.mz;
                for (int i = t0; i < cnt; i += nt) {
                    const Ix2 zr = split2(i, xh * ny, lxh + lny);
                    const Ix2 yx = split2(zr.lo, xh, lxh);
                    const int z = 2 + zr.hi, y = 2 + yx.hi;
                    const int x = 2 + ((par + y + z) & 1) + 2 * yx.lo;
                    if (x > g.mx || y > g.my || z > g.mz) continue;
                    rb3_point_pre(F.phi, F.rhs, F.cx, F.cy, F.cz, F.cc, F.at(z, y, x), F.sy, F.sz);
                }
            } else {
                const int cnt = xh * ny;
                for (int i = t0; i < cnt; i += nt) {
                    const Ix2 yx = split2(i, xh, lxh);
                    const int y = 2 + yx.hi;
                    const int x = 2 + ((par + y) & 1) + 2 * yx.lo;
                    if (x > g.mx || y > g.my) continue;
                    rb2_point_pre(F.phi, F.rhs, F.cx, F.cy, F.cc, F.at(y, x), F.sy);
                }
            }
        };
        if (op.kind == CO_PRO_RB) {
            // 2D only (the program builder fuses in 2D without D_xy): the
            // prolongation inside the half of parity a, k_rb2_pro's arithmetic.
            // This half reads only the other parity, each value taken as stored
            // value plus its correction, and nothing reads the other parity's
            // corrected values later (its half overwrites them without reading
            // them).
            if constexpr (dim == 2) {
                const LvPtr<T>& LC = lv[op.level + 1];
                const Geo gc = LC.g;
                const LvView<T, CL> C(LC, sm, !CL ? (long)b * gc.sz : 0, dim);
                const int par = op.a;
                auto C2 = [&](int J, int K) -> T { return C.phi[C.at(J, K)]; };
                auto nbv = [&](int xx, int yy, long q) -> T {
                    T v = F.phi[q];
                    T corr;
                    if (xx >= 2 && xx <= g.mx && yy >= 2 && yy <= g.my && prolong2_corr_at<T>(C2, xx, yy, 0, corr))
                        v = v + corr;
                    return v;
                };
                const int cnt = xh * ny;
                for (int i = t0; i < cnt; i += nt) {
                    const Ix2 yx = split2(i, xh, lxh);
                    const int y = 2 + yx.hi;
                    const int x = 2 + ((par + y) & 1) + 2 * yx.lo;
                    if (x > g.mx || y > g.my) continue;
                    const long p = F.at(y, x);
                    T acc = F.cy[p] * nbv(x, y + 1, p + F.sy) + F.cy[p - F.sy] * nbv(x, y - 1, p - F.sy);
                    acc = acc + F.cx[p] * nbv(x + 1, y, p + 1);
                    acc = acc + F.cx[p - 1] * nbv(x - 1, y, p - 1);
                    acc = acc - F.rhs[p];
                    F.phi[p] = acc / F.cc[p];
                }
            }
        } else {
            half(op.a);
        }
        if (op.kind == CO_RB_PAIR || (op.kind == CO_PRO_RB && op.b)) {
            // one wave (runs only): the second half reads what other lanes wrote
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            half(1 - op.a);
        }
    } else if (op.kind == CO_RB_DIAG) {
        const int xh = g.mx / 2, yh = g.my / 2, lxh = lg(xh);
        for (int i = t0; i < xh * yh; i += nt) {
            const Ix2 yx = split2(i, xh, lxh);
            const int y = (op.b ? 2 : 3) + 2 * yx.hi;
            const int x = (op.a ? 2 : 3) + 2 * yx.lo;
            if (y > g.my || x > g.mx) continue;
            rb2diag_point(F.phi, F.rhs, F.cx, F.cy, F.cc, F.drr, F.dlr, F.at(y, x), F.sy);
        }
    } else if (op.kind == CO_RESTRICT) {
        const LvPtr<T>& LC = lv[op.level + 1];
        const Geo gc = LC.g;
        const LvView<T, CL> C(LC, sm, (dim == 2 && !CL) ? (long)b * gc.sz : 0, dim);
        const int nx = gc.mx, ny = gc.my, lnx = lg(nx), lny = lg(ny);
        if constexpr (dim == 3) {
            // Four lanes per coarse point (a wave holds whole groups): part p
            // evaluates residuals 4p .. 4p+3 of restrict3_at's list, part 3 the
            // centre; the group's first lane folds them in restrict3_at's order,
            // so the value is restrict3_at's to the bit.
            const int cnt = nx * ny * gc.mz;
            const int part = t0 & 3, g0 = (t0 & (COARSE_WAVE - 1)) & ~3;
            const long A = part == 0 ? F.sy : F.sz, B = part == 2 ? F.sy : 1;
            for (int i4 = t0; i4 < 4 * cnt; i4 += nt) {
                const int i = i4 >> 2;
                const Ix2 zr = split2(i, nx * ny, lnx + lny);
                const Ix2 yx = split2(zr.lo, nx, lnx);
                const int I = 2 + zr.hi, J = 2 + yx.hi, K = 2 + yx.lo;
                const bool ok = K <= gc.mx && J <= gc.my && I <= gc.mz;
                T r0 = (T)0, r1 = (T)0, r2 = (T)0, r3 = (T)0;
                if (ok) {
                    const long c = F.at(2 * I - 1, 2 * J - 1, 2 * K - 1);
                    auto R = [&](long o) { return resid3<T, T>(F.phi, F.rhs, F.cx, F.cy, F.cz, F.cc, c + o, F.sy, F.sz); };
                    if (part < 3) {
                        r0 = R(A + B);
                        r1 = R(A - B);
                        r2 = R(-A + B);
                        r3 = R(-A - B);
                    } else {
                        r0 = R(0);
                    }
                }
                const T a0 = __shfl(r0, g0 + 1), a1 = __shfl(r1, g0 + 1), a2 = __shfl(r2, g0 + 1),
                        a3 = __shfl(r3, g0 + 1);
                const T b0 = __shfl(r0, g0 + 2), b1 = __shfl(r1, g0 + 2), b2 = __shfl(r2, g0 + 2),
                        b3 = __shfl(r3, g0 + 2);
                const T rcen = __shfl(r0, g0 + 3);
                if (ok && part == 0) {
                    T e = r0 + r1;
                    e = e + r2;
                    e = e + r3;
                    e = e + a0;
                    e = e + a1;
                    e = e + a2;
                    e = e + a3;
                    e = e + b0;
                    e = e + b1;
                    e = e + b2;
                    e = e + b3;
                    C.rhs[C.at(I, J, K)] = (T)0.03125 * e + (T)0.125 * rcen;
                }
            }
        } else {
            for (int i = t0; i < nx * ny; i += nt) {
                const Ix2 yx = split2(i, nx, lnx);
                const int J = 2 + yx.hi, K = 2 + yx.lo;
                if (K > gc.mx || J > gc.my) continue;
                const long c = F.at(2 * J - 1, 2 * K - 1);
                const T v = F.od ? restrict2_at(F.phi, F.rhs, F.cx, F.cy, F.cc, F.drr, F.dlr, c, F.sy)
                                 : restrict2_at_pre(F.phi, F.rhs, F.cx, F.cy, F.cc, c, F.sy);
                const long q = C.at(J, K);
                C.rhs[q] = v;
                if (op.a)
                    C.phi[q] = ((J + K) & 1) == 0 ? rb2_from_zero(v, RedZ<T>{C.cx, C.cy, nullptr, C.cc}, q, C.sy) : (T)0;
            }
        }
        // with a, every unknown of level + 1 was written above; its rim is
        // never written by any op and holds zeros
        if (op.b && !op.a)
            for (int i = t0; i < C.zn; i += nt) C.phi[i] = (T)0;
    } else if (op.kind == CO_ZERO) {
        for (int i = t0; i < F.zn; i += nt) F.phi[i] = (T)0;
    } else {  // CO_PROLONG
        const LvPtr<T>& LC = lv[op.level + 1];
        const Geo gc = LC.g;
        const LvView<T, CL> C(LC, sm, (dim == 2 && !CL) ? (long)b * gc.sz : 0, dim);
        const int nx = g.mx, ny = g.my, lnx = lg(nx), lny = lg(ny);
        if constexpr (dim == 3) {
            const int cnt = nx * ny * g.mz;
            for (int i = t0; i < cnt; i += nt) {
                const Ix2 zr = split2(i, nx * ny, lnx + lny);
                const Ix2 yx = split2(zr.lo, nx, lnx);
                const int z = 2 + zr.hi, y = 2 + yx.hi, x = 2 + yx.lo;
                if (x > g.mx || y > g.my || z > g.mz) continue;
                T corr;
                if (!prolong3_corr_at<T>([&](int I, int J, int K) -> T { return C.phi[C.at(I, J, K)]; }, x, y, z,
                                         op.a, corr))
                    continue;
                const long p = F.at(z, y, x);
                F.phi[p] = F.phi[p] + corr;
            }
        } else {
            for (int i = t0; i < nx * ny; i += nt) {
                const Ix2 yx = split2(i, nx, lnx);
                const int y = 2 + yx.hi, x = 2 + yx.lo;
                if (x > g.mx || y > g.my) continue;
                T corr;
                if (!prolong2_corr_at<T>([&](int J, int K) -> T { return C.phi[C.at(J, K)]; }, x, y, op.a, corr))
                    continue;
                const long p = F.at(y, x);
                F.phi[p] = F.phi[p] + corr;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Register runs (2D without D_xy): a run of one-wave ops on levels of at most
// 64 points each (padded: indices 2..m+1 per axis) keeps every point of up to
// three consecutive levels in a lane's registers, lane = yi * mx + xi for the
// point (2 + yi, 2 + xi); the last column and row are the boundary, whose phi
// stays zero.  Neighbours arrive from other lanes by shuffles instead of LDS
// round trips.  The arithmetic is rb2_point_pre / restrict2_at_pre /
// prolong2_corr_at's, in the same order, so the bits are the LDS ops'.
// ---------------------------------------------------------------------------
template <typename T>
struct RegLv {
    T phi, rhs, cxp, cxm, cyp, cym, cc;
    int mx, my, lmx;
};

template <typename T>
__device__ __forceinline__ void reg_load(RegLv<T>& S, const LvPtr<T>& L, T* sm, int lane)
{
    S.mx = L.g.mx;
    S.my = L.g.my;
    S.lmx = __builtin_ctz((unsigned)S.mx);
    const LvView<T, true> V(L, sm, 0, 2);
    const bool in = lane < S.mx * S.my;
    const int xi = lane & (S.mx - 1), yi = lane >> S.lmx;
    const long p = in ? V.at(2 + yi, 2 + xi) : V.at(2, 2);
    const T a0 = V.phi[p], a1 = V.rhs[p], a2 = V.cx[p], a3 = V.cx[p - 1], a4 = V.cy[p], a5 = V.cy[p - V.sy],
            a6 = V.cc[p];
    S.phi = in ? a0 : (T)0;
    S.rhs = a1;
    S.cxp = a2;
    S.cxm = a3;
    S.cyp = a4;
    S.cym = a5;
    S.cc = a6;
}

template <typename T>
__device__ __forceinline__ void reg_store(const RegLv<T>& S, const LvPtr<T>& L, T* sm, int lane)
{
    const LvView<T, true> V(L, sm, 0, 2);
    if (lane < S.mx * S.my) {
        const long p = V.at(2 + (lane >> S.lmx), 2 + (lane & (S.mx - 1)));
        V.phi[p] = S.phi;
        V.rhs[p] = S.rhs;
    }
}

// v at the four neighbours of this lane's point (v: the level's phi, or phi
// with the prolongation's corrections); 0 past the low boundary (index 1),
// the high boundary lanes hold 0
template <typename T>
__device__ __forceinline__ void reg_nbrs(const RegLv<T>& S, T v, int lane, T& pxp, T& pxm, T& pyp, T& pym)
{
    const int xi = lane & (S.mx - 1), yi = lane >> S.lmx;
    pxp = __shfl(v, (lane + 1) & 63, 64);
    pxm = __shfl(v, (lane - 1) & 63, 64);
    pyp = __shfl(v, (lane + S.mx) & 63, 64);
    pym = __shfl(v, (lane - S.mx) & 63, 64);
    if (xi == 0) pxm = (T)0;
    if (yi == 0) pym = (T)0;
}
template <typename T>
__device__ __forceinline__ void reg_nbrs(const RegLv<T>& S, int lane, T& pxp, T& pxm, T& pyp, T& pym)
{
    reg_nbrs(S, S.phi, lane, pxp, pxm, pyp, pym);
}

// one RBGS half of parity par; nb: the values its neighbours are read from
template <typename T>
__device__ __forceinline__ void reg_rb_from(RegLv<T>& S, T nb, int par, int lane)
{
    const int xi = lane & (S.mx - 1), yi = lane >> S.lmx;
    const bool upd = lane < S.mx * S.my && xi <= S.mx - 2 && yi <= S.my - 2 && ((xi + yi) & 1) == par;
    T pxp, pxm, pyp, pym;
    reg_nbrs(S, nb, lane, pxp, pxm, pyp, pym);
    T acc = S.cyp * pyp + S.cym * pym;
    acc = acc + S.cxp * pxp;
    acc = acc + S.cxm * pxm;
    acc = acc - S.rhs;
    const T v = acc / S.cc;
    S.phi = upd ? v : S.phi;
}

template <typename T>
__device__ __forceinline__ void reg_rb(RegLv<T>& S, int par, int lane)
{
    reg_rb_from(S, S.phi, par, lane);
}

// red: the coarse level's first half from phi = 0 as well (rb2_from_zero)
template <typename T>
__device__ __forceinline__ void reg_restrict(const RegLv<T>& F, RegLv<T>& C, int zero, int red, int lane)
{
    T pxp, pxm, pyp, pym;
    reg_nbrs(F, lane, pxp, pxm, pyp, pym);
    T r = F.cc * F.phi;
    r = r - F.cxm * pxm;
    r = r - F.cxp * pxp;
    r = r - F.cym * pym;
    r = r - F.cyp * pyp;
    r = r + F.rhs;
    // coarse point (2 + yc, 2 + xc) weights the fine points around fine lane (2yc + 1, 2xc + 1)
    const int xc = lane & (C.mx - 1), yc = lane >> C.lmx;
    const int fx = 2 * xc + 1, fy = 2 * yc + 1;
    auto R = [&](int yy, int xx) { return __shfl(r, (yy * F.mx + xx) & 63, 64); };
    const T ra = R(fy + 1, fx + 1), rb = R(fy + 1, fx - 1), rd = R(fy - 1, fx + 1), re = R(fy - 1, fx - 1),
            rc = R(fy, fx);
    T sum = ra + rb;
    sum = sum + rd;
    sum = sum + re;
    const T v = (T)0.0625 * sum + (T)0.25 * rc;
    const bool cunk = lane < C.mx * C.my && xc <= C.mx - 2 && yc <= C.my - 2;
    C.rhs = cunk ? v : C.rhs;
    if (red) {
        const T o = (T)0;
        T acc = C.cyp * o + C.cym * o;
        acc = acc + C.cxp * o;
        acc = acc + C.cxm * o;
        acc = acc - v;
        const T pv = acc / C.cc;
        C.phi = (cunk && ((xc + yc) & 1) == 0) ? pv : (T)0;
    } else {
        C.phi = zero ? (T)0 : C.phi;
    }
}

// the prolongation's correction of this lane's fine point; false: none
template <typename T>
__device__ __forceinline__ bool reg_corr(const RegLv<T>& F, const RegLv<T>& C, int full, int lane, T& corr)
{
    const int xi = lane & (F.mx - 1), yi = lane >> F.lmx;
    const int x = 2 + xi, y = 2 + yi;
    const int Xl = x >> 1, Yl = y >> 1;
    // coarse phi at coarse point (J, K); index 1 is the boundary (zero)
    auto Cv = [&](int J, int K) -> T {
        const bool ok = J >= 2 && K >= 2;
        const T v = __shfl(C.phi, ok ? ((J - 2) * C.mx + (K - 2)) & 63 : 0, 64);
        return ok ? v : (T)0;
    };
    const T A = Cv(Yl + 1, Xl + 1), B = Cv(Yl, Xl + 1), Cc = Cv(Yl + 1, Xl), D = Cv(Yl, Xl);
    const int ox = x & 1, oy = y & 1;
    // every case's value, then a select (no divergent branches)
    const T c_y = (T)0.5 * (A + Cc), c_x = (T)0.5 * (A + B);
    T sum = A + B;
    sum = sum + Cc;
    sum = sum + D;
    const T c_e = (T)0.25 * sum;
    corr = (ox && oy) ? A : oy ? c_y : ox ? c_x : c_e;
    return lane < F.mx * F.my && xi <= F.mx - 2 && yi <= F.my - 2 && (ox || oy || full);
}

// the prolongation (full = 0) inside the half of parity par, then (pair) the
// other half: the half reads its neighbours as phi plus their corrections
// (reg_prolong's values), and the other parity keeps its uncorrected phi,
// which its own half overwrites without reading
template <typename T>
__device__ __forceinline__ void reg_pro_rb(RegLv<T>& F, const RegLv<T>& C, int par, int pair, int lane)
{
    T corr;
    const bool has = reg_corr(F, C, 0, lane, corr);
    const T nb = has ? F.phi + corr : F.phi;
    reg_rb_from(F, nb, par, lane);
    if (pair) reg_rb(F, 1 - par, lane);
}

template <typename T>
__device__ __forceinline__ void reg_prolong(RegLv<T>& F, const RegLv<T>& C, int full, int lane)
{
    T corr;
    const bool upd = reg_corr(F, C, full, lane, corr);
    F.phi = upd ? F.phi + corr : F.phi;
}

// 2D levels of 65 to 256 points (padded indices 2..m+1) in a register run:
// lane (by, bx) = by * (mx / 2) + bx holds the 2 x 2 block of points
// (2 + 2by + dy, 2 + 2bx + dx), element k = 2 dy + dx.  The next coarser
// level's one-point layout (lane = yc * (mx / 2) + xc) puts coarse point
// (2 + yc, 2 + xc) in the lane of the block that holds its restriction centre
// (element 3) and one of its corners (element 0), so the restriction and the
// prolongation between the two levels need only three shuffles.
template <typename T>
struct RegLv4 {
    T phi[4], rhs[4], cxp[4], cyp[4], cc[4];
    // cx[p - 1] of elements 0 and 2, cy[p - sy] of elements 0 and 1; those of
    // the other elements are the lane's own cxp[0], cxp[2] and cyp[0], cyp[1]
    T cxm0, cxm2, cym0, cym1;
    int w, h, lw;   // blocks per row (mx / 2), block rows (my / 2), log2(w)
};

template <typename T>
__device__ __forceinline__ void reg4_load(RegLv4<T>& Q, const LvPtr<T>& L, T* sm, int lane)
{
    Q.w = L.g.mx >> 1;
    Q.h = L.g.my >> 1;
    Q.lw = __builtin_ctz((unsigned)Q.w);
    const LvView<T, true> V(L, sm, 0, 2);
    const bool in = lane < Q.w * Q.h;
    const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
    const int y0 = in ? 2 + 2 * by : 2, x0 = in ? 2 + 2 * bx : 2;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long p = V.at(y0 + (k >> 1), x0 + (k & 1));
        const T a = V.phi[p];
        Q.phi[k] = in ? a : (T)0;
        Q.rhs[k] = V.rhs[p];
        Q.cxp[k] = V.cx[p];
        Q.cyp[k] = V.cy[p];
        Q.cc[k] = V.cc[p];
    }
    const long p0 = V.at(y0, x0), p1 = V.at(y0, x0 + 1), p2 = V.at(y0 + 1, x0);
    Q.cxm0 = V.cx[p0 - 1];
    Q.cxm2 = V.cx[p2 - 1];
    Q.cym0 = V.cy[p0 - V.sy];
    Q.cym1 = V.cy[p1 - V.sy];
}

template <typename T>
__device__ __forceinline__ void reg4_store(const RegLv4<T>& Q, const LvPtr<T>& L, T* sm, int lane)
{
    const LvView<T, true> V(L, sm, 0, 2);
    if (lane < Q.w * Q.h) {
        const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long p = V.at(2 + 2 * by + (k >> 1), 2 + 2 * bx + (k & 1));
            V.phi[p] = Q.phi[k];
            V.rhs[p] = Q.rhs[k];
        }
    }
}

// element k of this lane's block is an unknown (not the boundary row / column)
template <typename T>
__device__ __forceinline__ bool reg4_unk(const RegLv4<T>& Q, int lane, int k)
{
    const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
    return lane < Q.w * Q.h && 2 * bx + (k & 1) <= 2 * Q.w - 2 && 2 * by + (k >> 1) <= 2 * Q.h - 2;
}

// one RBGS half of parity par (elements 0 and 3 for par 0, 1 and 2 for par
// 1), its neighbours read from n0..n3 (phi, or phi plus the prolongation's
// corrections); rb2_point_pre's arithmetic in its order
template <typename T>
__device__ __forceinline__ void reg4_rb_from(RegLv4<T>& Q, T n0, T n1, T n2, T n3, int par, int lane)
{
    const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
    auto sh = [&](T v, int src) { return __shfl(v, src & 63, 64); };
    if (par == 0) {
        T xm0 = sh(n1, lane - 1), ym0 = sh(n2, lane - Q.w);
        const T xp3 = sh(n2, lane + 1), yp3 = sh(n1, lane + Q.w);
        if (bx == 0) xm0 = (T)0;
        if (by == 0) ym0 = (T)0;
        T a0 = Q.cyp[0] * n2 + Q.cym0 * ym0;
        a0 = a0 + Q.cxp[0] * n1;
        a0 = a0 + Q.cxm0 * xm0;
        a0 = a0 - Q.rhs[0];
        T a3 = Q.cyp[3] * yp3 + Q.cyp[1] * n1;
        a3 = a3 + Q.cxp[3] * xp3;
        a3 = a3 + Q.cxp[2] * n2;
        a3 = a3 - Q.rhs[3];
        const T v0 = a0 / Q.cc[0], v3 = a3 / Q.cc[3];
        Q.phi[0] = reg4_unk(Q, lane, 0) ? v0 : Q.phi[0];
        Q.phi[3] = reg4_unk(Q, lane, 3) ? v3 : Q.phi[3];
    } else {
        T ym1 = sh(n3, lane - Q.w), xm2 = sh(n3, lane - 1);
        const T xp1 = sh(n0, lane + 1), yp2 = sh(n0, lane + Q.w);
        if (by == 0) ym1 = (T)0;
        if (bx == 0) xm2 = (T)0;
        T a1 = Q.cyp[1] * n3 + Q.cym1 * ym1;
        a1 = a1 + Q.cxp[1] * xp1;
        a1 = a1 + Q.cxp[0] * n0;
        a1 = a1 - Q.rhs[1];
        T a2 = Q.cyp[2] * yp2 + Q.cyp[0] * n0;
        a2 = a2 + Q.cxp[2] * n3;
        a2 = a2 + Q.cxm2 * xm2;
        a2 = a2 - Q.rhs[2];
        const T v1 = a1 / Q.cc[1], v2 = a2 / Q.cc[2];
        Q.phi[1] = reg4_unk(Q, lane, 1) ? v1 : Q.phi[1];
        Q.phi[2] = reg4_unk(Q, lane, 2) ? v2 : Q.phi[2];
    }
}

template <typename T>
__device__ __forceinline__ void reg4_rb(RegLv4<T>& Q, int par, int lane)
{
    reg4_rb_from(Q, Q.phi[0], Q.phi[1], Q.phi[2], Q.phi[3], par, lane);
}

// restriction of the 4-point level Q into the one-point level C
// (reg_restrict's arithmetic: residuals in resid2's order at the corner
// element 0 and the centre element 3, weighted in restrict2_at_pre's order)
template <typename T>
__device__ __forceinline__ void reg4_restrict(const RegLv4<T>& Q, RegLv<T>& C, int zero, int red, int lane)
{
    const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
    auto sh = [&](T v, int src) { return __shfl(v, src & 63, 64); };
    T xm0 = sh(Q.phi[1], lane - 1), ym0 = sh(Q.phi[2], lane - Q.w);
    const T xp3 = sh(Q.phi[2], lane + 1), yp3 = sh(Q.phi[1], lane + Q.w);
    if (bx == 0) xm0 = (T)0;
    if (by == 0) ym0 = (T)0;
    T r0 = Q.cc[0] * Q.phi[0];
    r0 = r0 - Q.cxm0 * xm0;
    r0 = r0 - Q.cxp[0] * Q.phi[1];
    r0 = r0 - Q.cym0 * ym0;
    r0 = r0 - Q.cyp[0] * Q.phi[2];
    r0 = r0 + Q.rhs[0];
    T r3 = Q.cc[3] * Q.phi[3];
    r3 = r3 - Q.cxp[2] * Q.phi[2];
    r3 = r3 - Q.cxp[3] * xp3;
    r3 = r3 - Q.cyp[1] * Q.phi[1];
    r3 = r3 - Q.cyp[3] * yp3;
    r3 = r3 + Q.rhs[3];
    // coarse point (2 + by, 2 + bx): corners (y + 1, x + 1), (y + 1, x - 1),
    // (y - 1, x + 1), (y - 1, x - 1) of its centre are element 0 of the blocks
    // (by + 1, bx + 1), (by + 1, bx), (by, bx + 1) and this one
    const T ra = sh(r0, lane + Q.w + 1), rb = sh(r0, lane + Q.w), rd = sh(r0, lane + 1);
    T sum = ra + rb;
    sum = sum + rd;
    sum = sum + r0;
    const T v = (T)0.0625 * sum + (T)0.25 * r3;
    const bool cunk = lane < C.mx * C.my && bx <= C.mx - 2 && by <= C.my - 2;
    C.rhs = cunk ? v : C.rhs;
    if (red) {
        const T o = (T)0;
        T acc = C.cyp * o + C.cym * o;
        acc = acc + C.cxp * o;
        acc = acc + C.cxm * o;
        acc = acc - v;
        const T pv = acc / C.cc;
        C.phi = (cunk && ((bx + by) & 1) == 0) ? pv : (T)0;
    } else {
        C.phi = zero ? (T)0 : C.phi;
    }
}

// the prolongation's corrections of the block's elements from the one-point
// level C (reg_corr's values); has[k]: element k takes its correction
template <typename T>
__device__ __forceinline__ void reg4_corr(const RegLv4<T>& Q, const RegLv<T>& C, int full, int lane, T (&corr)[4],
                                          bool (&has)[4])
{
    const int bx = lane & (Q.w - 1), by = lane >> Q.lw;
    auto sh = [&](T v, int src) { return __shfl(v, src & 63, 64); };
    // coarse (2 + by, 2 + bx), (1 + by, 2 + bx), (2 + by, 1 + bx), (1 + by, 1 + bx); index 1 is the boundary
    const T A = C.phi;
    T B = sh(C.phi, lane - Q.w), Cc = sh(C.phi, lane - 1), D = sh(C.phi, lane - Q.w - 1);
    if (by == 0) {
        B = (T)0;
        D = (T)0;
    }
    if (bx == 0) {
        Cc = (T)0;
        D = (T)0;
    }
    const T c_y = (T)0.5 * (A + Cc), c_x = (T)0.5 * (A + B);
    T sum = A + B;
    sum = sum + Cc;
    sum = sum + D;
    corr[0] = (T)0.25 * sum;   // (even, even)
    corr[1] = c_x;             // x odd
    corr[2] = c_y;             // y odd
    corr[3] = A;               // (odd, odd)
#pragma unroll
    for (int k = 0; k < 4; ++k) has[k] = reg4_unk(Q, lane, k) && (k != 0 || full);
}

template <typename T>
__device__ __forceinline__ void reg4_prolong(RegLv4<T>& Q, const RegLv<T>& C, int full, int lane)
{
    T corr[4];
    bool has[4];
    reg4_corr(Q, C, full, lane, corr, has);
#pragma unroll
    for (int k = 0; k < 4; ++k) Q.phi[k] = has[k] ? Q.phi[k] + corr[k] : Q.phi[k];
}

// the prolongation (full = 0) inside the half of parity par, then (pair) the
// other half, as reg_pro_rb
template <typename T>
__device__ __forceinline__ void reg4_pro_rb(RegLv4<T>& Q, const RegLv<T>& C, int par, int pair, int lane)
{
    T corr[4];
    bool has[4];
    reg4_corr(Q, C, 0, lane, corr, has);
    T nb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nb[k] = has[k] ? Q.phi[k] + corr[k] : Q.phi[k];
    reg4_rb_from(Q, nb[0], nb[1], nb[2], nb[3], par, lane);
    if (pair) reg4_rb(Q, 1 - par, lane);
}

// A run whose finest level lmin holds 65 to 256 points: that level in the
// 4-point layout, up to three coarser one-point levels.
template <typename T>
__device__ __forceinline__ void coarse_reg_run4(const unsigned* sops, int o, int run, int lmin, int lmax,
                                                const LvPtr<T>* __restrict__ lv, T* sm, int lane)
{
    RegLv4<T> Q;
    RegLv<T> S1, S2, S3;
    reg4_load(Q, lv[lmin], sm, lane);
    if (lmin + 1 <= lmax) reg_load(S1, lv[lmin + 1], sm, lane);
    if (lmin + 2 <= lmax) reg_load(S2, lv[lmin + 2], sm, lane);
    if (lmin + 3 <= lmax) reg_load(S3, lv[lmin + 3], sm, lane);
    unsigned wn = sops[o];
    for (int k = 0; k < run; ++k) {
        const CoarseOp q = coarse_unpack(__builtin_amdgcn_readfirstlane(wn));
        if (k + 1 < run) wn = sops[o + k + 1];   // next op word, read ahead
        const int j = q.level - lmin;
        if (j == 0) {
            if (q.kind == CO_RB || q.kind == CO_RB_PAIR) {
                reg4_rb(Q, q.a, lane);
                if (q.kind == CO_RB_PAIR) reg4_rb(Q, 1 - q.a, lane);
            } else if (q.kind == CO_RESTRICT) {
                reg4_restrict(Q, S1, q.b, q.a, lane);
            } else if (q.kind == CO_PRO_RB) {
                reg4_pro_rb(Q, S1, q.a, q.b, lane);
            } else if (q.kind == CO_PROLONG) {
                reg4_prolong(Q, S1, q.a, lane);
            } else {  // CO_ZERO
#pragma unroll
                for (int e = 0; e < 4; ++e) Q.phi[e] = (T)0;
            }
            continue;
        }
        if (q.kind == CO_RB || q.kind == CO_RB_PAIR) {
            auto sweep = [&](RegLv<T>& S) {
                reg_rb(S, q.a, lane);
                if (q.kind == CO_RB_PAIR) reg_rb(S, 1 - q.a, lane);
            };
            if (j == 1) sweep(S1);
            else if (j == 2) sweep(S2);
            else sweep(S3);
        } else if (q.kind == CO_RESTRICT) {
            if (j == 1) reg_restrict(S1, S2, q.b, q.a, lane);
            else reg_restrict(S2, S3, q.b, q.a, lane);
        } else if (q.kind == CO_PRO_RB) {
            if (j == 1) reg_pro_rb(S1, S2, q.a, q.b, lane);
            else reg_pro_rb(S2, S3, q.a, q.b, lane);
        } else if (q.kind == CO_PROLONG) {
            if (j == 1) reg_prolong(S1, S2, q.a, lane);
            else reg_prolong(S2, S3, q.a, lane);
        } else {  // CO_ZERO
            if (j == 1) S1.phi = (T)0;
            else if (j == 2) S2.phi = (T)0;
            else S3.phi = (T)0;
        }
    }
    reg4_store(Q, lv[lmin], sm, lane);
    if (lmin + 1 <= lmax) reg_store(S1, lv[lmin + 1], sm, lane);
    if (lmin + 2 <= lmax) reg_store(S2, lv[lmin + 2], sm, lane);
    if (lmin + 3 <= lmax) reg_store(S3, lv[lmin + 3], sm, lane);
}

// The run of `run` ops from word o, in wave 0: the levels it touches (at most
// three of one point per lane, lmin..lmax, or a 4-point level and three
// more) are loaded into registers, the ops execute there, and phi and rhs go
// back to LDS.
template <typename T>
__device__ __forceinline__ void coarse_reg_run(const unsigned* sops, int o, int run, const LvPtr<T>* __restrict__ lv,
                                               T* sm, int lane)
{
    int lmin = 31, lmax = 0;
    for (int k = 0; k < run; ++k) {
        const CoarseOp q = coarse_unpack(__builtin_amdgcn_readfirstlane(sops[o + k]));
        const int hi = (q.kind == CO_RESTRICT || q.kind == CO_PROLONG || q.kind == CO_PRO_RB) ? q.level + 1 : q.level;
        lmin = q.level < lmin ? q.level : lmin;
        lmax = hi > lmax ? hi : lmax;
    }
    // a finest level of more than one point per lane: the 4-point layout
    if ((long)lv[lmin].g.mx * lv[lmin].g.my > COARSE_WAVE) {
        coarse_reg_run4<T>(sops, o, run, lmin, lmax, lv, sm, lane);
        return;
    }
    RegLv<T> S0, S1, S2;
    reg_load(S0, lv[lmin], sm, lane);
    if (lmin + 1 <= lmax) reg_load(S1, lv[lmin + 1], sm, lane);
    if (lmin + 2 <= lmax) reg_load(S2, lv[lmin + 2], sm, lane);
    unsigned wn = sops[o];
    for (int k = 0; k < run; ++k) {
        const CoarseOp q = coarse_unpack(__builtin_amdgcn_readfirstlane(wn));
        if (k + 1 < run) wn = sops[o + k + 1];   // next op word, read ahead
        const int j = q.level - lmin;
        if (q.kind == CO_RB || q.kind == CO_RB_PAIR) {
            auto sweep = [&](RegLv<T>& S) {
                reg_rb(S, q.a, lane);
                if (q.kind == CO_RB_PAIR) reg_rb(S, 1 - q.a, lane);
            };
            if (j == 0) sweep(S0);
            else if (j == 1) sweep(S1);
            else sweep(S2);
        } else if (q.kind == CO_RESTRICT) {
            if (j == 0) reg_restrict(S0, S1, q.b, q.a, lane);
            else reg_restrict(S1, S2, q.b, q.a, lane);
        } else if (q.kind == CO_PRO_RB) {
            if (j == 0) reg_pro_rb(S0, S1, q.a, q.b, lane);
            else reg_pro_rb(S1, S2, q.a, q.b, lane);
        } else if (q.kind == CO_PROLONG) {
            if (j == 0) reg_prolong(S0, S1, q.a, lane);
            else reg_prolong(S1, S2, q.a, lane);
        } else {  // CO_ZERO
            if (j == 0) S0.phi = (T)0;
            else if (j == 1) S1.phi = (T)0;
            else S2.phi = (T)0;
        }
    }
    reg_store(S0, lv[lmin], sm, lane);
    if (lmin + 1 <= lmax) reg_store(S1, lv[lmin + 1], sm, lane);
    if (lmin + 2 <= lmax) reg_store(S2, lv[lmin + 2], sm, lane);
}

// Coordinates (z, y, x) of linear index i over planes of ny rows of nx,
// advanced by a fixed stride: one division at the start, none per step.
struct Walk3 {
    int z, y, x, dz, dy, dx, ny, nx;
    __device__ __forceinline__ Walk3(int i, int stride, int ny_, int nx_) : ny(ny_), nx(nx_)
    {
        const int nxy = ny * nx;
        z = i / nxy;
        const int r = i - z * nxy;
        y = r / nx;
        x = r - y * nx;
        dz = stride / nxy;
        const int q = stride - dz * nxy;
        dy = q / nx;
        dx = q - dy * nx;
    }
    __device__ __forceinline__ void next()
    {
        x += dx;
        y += dy;
        z += dz;
        if (x >= nx) {
            x -= nx;
            ++y;
        }
        if (y >= ny) {
            y -= ny;
            ++z;
        }
    }
};

// ops: nops packed op words, copied to LDS after the levels (element offset
// ops_at of T); img: this launch's image of the staged levels (img_n
// elements per problem, a multiple of 16 bytes; problem stride img_bs),
// copied to LDS offset 0; levels
// [lds_from, lds_to) are in LDS; wb_level (-1: none) is the start level when
// it is staged.
// One workgroup per CU: at most NT / 256 waves per SIMD, which lets the
// scheduler spend registers on latency (loads issued early) instead of on an
// occupancy the launch never reaches.
template <typename T, int NT, int dim>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, NT / 256)))
k_coarse_prog(const LvPtr<T>* __restrict__ lv, const unsigned* __restrict__ ops, int nops, int ops_at,
              const int* __restrict__ act, int lds_from, int lds_to, int wb_level, const T* __restrict__ img,
              int img_n, long img_bs)
{
    const int b = blockIdx.x;               // 2D: problem of the batch; 3D: 0
    if (act != nullptr && act[b] == 0) return;
    const int tid = threadIdx.x;
    T* sm = reinterpret_cast<T*>(jsv_dyn_smem);
    unsigned* sops = reinterpret_cast<unsigned*>(sm + ops_at);
    {
        // The image goes global -> LDS directly (16 bytes per lane, no
        // registers; wave w moves pieces [c, c + 64) to LDS piece c); the op
        // words and the start level's rhs (after the image: no overlap) go
        // through registers with all their loads issued before any store.
        const int n16 = (int)(img_n * sizeof(T) / 16);
        const unsigned* isrc = reinterpret_cast<const unsigned*>(img + (long)b * img_bs);
        const int lane = tid & (COARSE_WAVE - 1);
        for (int c = tid - lane; c < n16; c += NT)
            if (c + lane < n16)
                __builtin_amdgcn_global_load_lds(isrc + 4 * (c + lane),
                                                 (__attribute__((address_space(3))) void*)(sm + c * (16 / sizeof(T))),
                                                 16, 0, 0);
        const LvPtr<T>* W = wb_level >= 0 ? &lv[wb_level] : nullptr;
        const int nr = W != nullptr ? W->ln : 0;
        const T* rsrc = W != nullptr ? W->rhs + (dim == 2 ? (long)b * W->g.sz : 0) : nullptr;
        T* rdst = W != nullptr ? sm + W->lds[1] : nullptr;
        constexpr int UO = 2, UR = 8;
        auto cdiv = [](int a, int q) { return (a + q - 1) / q; };
        const int rounds = max(cdiv(nops, NT * UO), cdiv(nr, NT * UR));
        const long rsz = W != nullptr ? W->g.sz : 0, rsy = W != nullptr ? W->g.sy : 0;
        Walk3 wr(tid, NT, W != nullptr ? W->lsz / W->lsy : 1, W != nullptr ? W->lsy : 1);
        for (int r = 0; r < rounds; ++r) {
            unsigned vo[UO];
            T vr[UR];
#pragma unroll
            for (int k = 0; k < UO; ++k) {
                const int i = (r * UO + k) * NT + tid;
                vo[k] = ops[i < nops ? i : nops - 1];
            }
            if (nr > 0) {
#pragma unroll
                for (int k = 0; k < UR; ++k) {
                    const int i = (r * UR + k) * NT + tid;
                    // compact point i is padded point (z, y, x) + 1; past the
                    // end the load reads point 0 (value unused)
                    const bool in = i < nr;
                    const int z = in ? wr.z : 0, y = in ? wr.y : 0, x = in ? wr.x : 0;
                    vr[k] = rsrc[(dim == 3 ? (long)(z + 1) * rsz : 0) + (long)(y + 1) * rsy + x + 1];
                    wr.next();
                }
            }
#pragma unroll
            for (int k = 0; k < UO; ++k) {
                const int i = (r * UO + k) * NT + tid;
                if (i < nops) sops[i] = vo[k];
            }
            if (nr > 0) {
#pragma unroll
                for (int k = 0; k < UR; ++k) {
                    const int i = (r * UR + k) * NT + tid;
                    if (i < nr) rdst[i] = vr[k];
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);   // the direct-to-LDS loads have landed
    }
    __syncthreads();
    for (int o = 0; o < nops;) {
        const CoarseOp op = coarse_unpack(__builtin_amdgcn_readfirstlane(sops[o]));
        if (op.run > 0 && dim == 2 && op.reg) {
            if (tid < COARSE_WAVE) coarse_reg_run<T>(sops, o, op.run, lv, sm, tid);
            // the ops after the run read what the wave stored
            o += op.run;
        } else if (op.run > 0) {
            if (tid < COARSE_WAVE) {
                unsigned wn = sops[o + 1 < nops ? o + 1 : o];   // next op word, read ahead
                for (int k = 0; k < op.run; ++k) {
                    const CoarseOp q = k == 0 ? op : coarse_unpack(__builtin_amdgcn_readfirstlane(wn));
                    if (k > 0) wn = sops[o + k + 1 < nops ? o + k + 1 : o + k];
                    coarse_op<T, dim, true, true>(q, lv, sm, b, tid, COARSE_WAVE);
                    // the next op reads what other lanes of this wave wrote: no
                    // reordering of LDS accesses across this point
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            o += op.run;
        } else {
            if (op.level >= lds_from && op.level < lds_to)
                coarse_op<T, dim, true, true>(op, lv, sm, b, tid, NT);
            else if (op.level + 1 >= lds_from && op.level + 1 < lds_to)
                coarse_op<T, dim, false, true>(op, lv, sm, b, tid, NT);
            else
                coarse_op<T, dim, false, false>(op, lv, sm, b, tid, NT);
            ++o;
        }
        __syncthreads();
    }
    if (wb_level >= 0) {
        // the staged start level's phi goes back, whole padded array (ghost layers 0)
        const LvPtr<T>& W = lv[wb_level];
        const int np = (int)(dim == 3 ? W.g.n : W.g.sz);
        T* dst = W.phi + (dim == 2 ? (long)b * W.g.sz : 0);
        const T* s = sm + W.lds[0];
        const int nx = W.g.nx, ny = W.g.ny;
        constexpr int U = 8;
        Walk3 w(tid, NT, ny, nx);
        for (int i0 = 0; i0 < np; i0 += NT * U) {
            T v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                // padded point (z, y, x); 2D: one plane (w.z == 0 inside the array)
                const int z = dim == 3 ? w.z : (w.z == 0 ? 1 : 0), y = w.y, x = w.x;
                const bool in = z >= 1 && y >= 1 && x >= 1 && y <= ny - 2 && x <= nx - 2 &&
                                (dim == 2 || z <= W.g.nz - 2);
                v[k] = in ? s[(dim == 3 ? (z - 1) * W.lsz : 0) + (y - 1) * W.lsy + x - 1] : (T)0;
                w.next();
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int i = i0 + k * NT + tid;
                if (i < np) dst[i] = v[k];
            }
        }
    }
}

// Image of the staged levels (the LDS region in its layout): one array of
// level geometry g gathered into its compact copy at offset dst of every
// problem's image (stride dbs); src has problem stride sbs.
template <typename T>
__global__ void __launch_bounds__(256)
k_stage_image(T* __restrict__ img, long dbs, int dst, const T* __restrict__ src, long sbs, Geo g, int lsy,
              int lsz, int ln, int dim)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ln) return;
    const long b = blockIdx.y;
    const int z = i / lsz, r = i - z * lsz, y = r / lsy, x = r - y * lsy;
    const long q = (dim == 3 ? (long)(z + 1) * g.sz : 0) + (long)(y + 1) * g.sy + x + 1;
    img[b * dbs + dst + i] = src[b * sbs + q];
}

// ---------------------------------------------------------------------------
// Reductions.  Wave64 shuffles, then LDS across the 4 waves of a workgroup.
// ---------------------------------------------------------------------------

// max |a - b| over n elements (distance, solver_2d.py:481-494).  Non-negative
// doubles order like their bit patterns, so a 64-bit atomicMax on the bits is
// exact and order-independent; a NaN (positive after fabs) compares above
// +inf, so a NaN anywhere yields NaN as jnp.max does.
template <typename T>
__global__ void __launch_bounds__(256)
k_maxabsdiff(const T* __restrict__ a0, const T* __restrict__ b0, long n, unsigned long long* out0)
{
    // blockIdx.y selects a segment of n elements (one problem of a batch)
    __shared__ unsigned long long sh[4];
    const T* __restrict__ a = a0 + (long)blockIdx.y * n;
    const T* __restrict__ b = b0 + (long)blockIdx.y * n;
    unsigned long long* out = out0 + blockIdx.y;
    unsigned long long best = 0ull;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double d = fabs((double)a[i] - (double)b[i]);
        unsigned long long u = (unsigned long long)__double_as_longlong(d);
        best = u > best ? u : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long v = __shfl_xor(best, o, 64);
        best = v > best ? v : best;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sh[0];
        for (int i = 1; i < 4; ++i) m = sh[i] > m ? sh[i] : m;
        atomicMax(out, m);
    }
}

// Residual rhs - L phi over the unknowns of the fine level, in the arithmetic
// type T of the operator; optionally stores it (as TO) and writes per-block
// partial sums of r^2 (FP64) for a deterministic two-pass norm.  act: one
// activity flag for the whole launch (nullptr: always runs).
template <typename T, typename TP, typename TO>
__global__ void __launch_bounds__(BX * BY)
k_residual(const TP* __restrict__ phi, const TP* __restrict__ rhs, const T* __restrict__ cx,
           const T* __restrict__ cy, const T* __restrict__ cz, const T* __restrict__ cc,
           const T* __restrict__ drr, const T* __restrict__ dlr, Geo g, TO* __restrict__ rout,
           double* __restrict__ partials, int dim, const int* __restrict__ act)
{
    if (act != nullptr && act[0] == 0) return;
    __shared__ double sh[BY];
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : (int)blockIdx.z;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    double v = 0.0;
    if (y <= g.my && x <= g.mx) {
        const long p = (long)z * g.sz + (long)y * g.sy + x;
        T r;
        if (dim == 3) r = resid3<T, TP>(phi, rhs, cx, cy, cz, cc, p, g.sy, g.sz);
        else          r = resid2<T, TP>(phi, rhs, cx, cy, cc, drr, dlr, p, g.sy);
        if (rout != nullptr) rout[p] = (TO)r;
        v = (double)r * (double)r;
    }
    v = wave_sum(v);
    if (threadIdx.x == 0) sh[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double s = sh[0];
        for (int i = 1; i < BY; ++i) s += sh[i];
        partials[((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
    }
}

// sum of squares over the unknowns of one array (||rhs||_2^2)
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_norm2_unknowns(const T* __restrict__ a, Geo g, double* __restrict__ partials, int dim)
{
    __shared__ double sh[BY];
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : (int)blockIdx.z;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    double v = 0.0;
    if (y <= g.my && x <= g.mx) {
        const double q = (double)a[(long)z * g.sz + (long)y * g.sy + x];
        v = q * q;
    }
    v = wave_sum(v);
    if (threadIdx.x == 0) sh[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double s = sh[0];
        for (int i = 1; i < BY; ++i) s += sh[i];
        partials[((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
    }
}

// Sum of seg[0, n) in a fixed order by a workgroup of 1024 threads; the
// result is valid in thread 0.
__device__ __forceinline__ double block_sum_fixed(const double* __restrict__ seg, long n, double* sh)
{
    double s = 0.0;
    // each thread adds its elements in increasing index order; eight loads are
    // issued before their adds, so the serial sum waits on one memory latency
    // per eight elements instead of one per element
    constexpr int U = 8;
    for (long i0 = threadIdx.x; i0 < n; i0 += 1024L * U) {
        double v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long i = i0 + 1024L * k;
            v[k] = i < n ? seg[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (i0 + 1024L * k < n) s += v[k];
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < 16; ++i) t += sh[i];
    return t;
}

// Second pass: workgroup g sums partials [g*n, (g+1)*n) in a fixed order
// (one segment per batched problem, or one for a whole 3D level).
__global__ void __launch_bounds__(1024)
k_sum_partials(const double* __restrict__ partials, long n, double* __restrict__ out)
{
    __shared__ double sh[16];
    const double t = block_sum_fixed(partials + (long)blockIdx.x * n, n, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Stop rule on the device (solver_2d.py:464-477).  A dynamic solve of one
// problem replays its cycles with the test inside: a one-workgroup kernel
// evaluates the stop quantity, counts the cycle and, on a stop, clears the
// activity flag that every kernel of the solve checks first, so the cycles
// queued behind it do nothing and the host reads the outcome once per block
// of cycles instead of after every cycle.
// ---------------------------------------------------------------------------
struct StopCtl {
    double thr;              // tolerance: rtol (relative residual) or epsilon (distance)
    double rhs_norm;         // ||rhs|| over the unknowns (relative residual)
    double last;             // stop value of the last tested cycle
    long long max_iters;     // cycle limit
    long long cycle;         // kernels of the test run so far (one per cycle)
    long long count;         // cycles tested
    int stopped;
    int nonfinite;           // a stop value was NaN or infinite
    double* hist;            // stop value of cycle k at hist[k - 1] (nullptr: not kept)
    long long hist_len;
};

// mode 0 (relative residual, solver_2d.py rel. residual rule): runs at the
//   start of a cycle; partials[0, n) hold the squared residual of the
//   iterate the cycle starts from, summed in block_sum_fixed's order, so the
//   value has the bits of the host's check.  The first cycle starts from the
//   initial guess, which is not tested.
// mode 1 (distance, solver_2d.py:481-494): runs at the end of a cycle;
//   partials[0, n) hold per-workgroup maxima of |phi - phi_old| as bit
//   patterns (k_stop_maxdiff).
// A cycle is the last when its value is not above the tolerance (NaN
// included: `dist > eps` is false) or the cycle limit is reached.
__global__ void __launch_bounds__(1024)
k_stop_check(const double* __restrict__ partials, long n, int mode, StopCtl* __restrict__ ctl,
             int* __restrict__ act)
{
    if (act[0] == 0) return;
    __shared__ double sh[16];
    __shared__ unsigned long long shm[16];
    double v = 0.0;
    if (mode == 0) {
        v = block_sum_fixed(partials, n, sh);
    } else {
        const unsigned long long* pu = reinterpret_cast<const unsigned long long*>(partials);
        unsigned long long m = 0ull;
        for (long i = threadIdx.x; i < n; i += 1024) m = pu[i] > m ? pu[i] : m;
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long u = __shfl_xor(m, o, 64);
            m = u > m ? u : m;
        }
        if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long mm = shm[0];
            for (int i = 1; i < 16; ++i) mm = shm[i] > mm ? shm[i] : mm;
            v = __longlong_as_double((long long)mm);
        }
    }
    if (threadIdx.x != 0) return;
    const long long c = ctl->cycle + 1;
    ctl->cycle = c;
    long long k = c;
    if (mode == 0) {
        if (c == 1) return;
        v = sqrt(v) / ctl->rhs_norm;
        k = c - 1;
    }
    ctl->last = v;
    ctl->count = k;
    if (!isfinite(v)) ctl->nonfinite = 1;
    if (ctl->hist != nullptr && k - 1 < ctl->hist_len) ctl->hist[k - 1] = v;
    if (!(v > ctl->thr) || k >= ctl->max_iters) {
        ctl->stopped = 1;
        act[0] = 0;
    }
}

// Per-workgroup max |a - b| over n elements, as the bit pattern of the
// non-negative double (see k_maxabsdiff), into part[blockIdx.x].
__global__ void __launch_bounds__(256)
k_stop_maxdiff(const double* __restrict__ a, const double* __restrict__ b, long n, double* __restrict__ part,
               const int* __restrict__ act)
{
    if (act[0] == 0) return;
    __shared__ unsigned long long sh[4];
    unsigned long long best = 0ull;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double d = fabs(a[i] - b[i]);
        const unsigned long long u = (unsigned long long)__double_as_longlong(d);
        best = u > best ? u : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long u = __shfl_xor(best, o, 64);
        best = u > best ? u : best;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sh[0];
        for (int i = 1; i < 4; ++i) m = sh[i] > m ? sh[i] : m;
        reinterpret_cast<unsigned long long*>(part)[blockIdx.x] = m;
    }
}

// dst := src (n elements) at the start of every cycle but the solve's first,
// whose phi_old is the caller's initial phi, put there by the host.
__global__ void __launch_bounds__(256)
k_stop_copy(double* __restrict__ dst, const double* __restrict__ src, long n, const StopCtl* __restrict__ ctl,
            const int* __restrict__ act)
{
    if (act[0] == 0 || ctl->cycle == 0) return;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ a, long n, T v)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) a[i] = v;
}

__global__ void __launch_bounds__(64) k_set_u64(unsigned long long* a, unsigned long long v)
{
    if (threadIdx.x == 0) a[0] = v;
}

// Streaming bandwidth reference: dst[i] = sum over the K source arrays
// (consecutive in src, n floats each) of src[a n + i], as 16-byte vectors.
// K = 1 is a copy; K = 6 reads and writes in the RBGS sweep's ratio (phi,
// rhs, three faces and the diagonal read; phi written).
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <int K>
__global__ void __launch_bounds__(256) k_stream(const f32x4* __restrict__ src, f32x4* __restrict__ dst, long n4)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long)gridDim.x * 256) {
        f32x4 v[K];
#pragma unroll
        for (int a = 0; a < K; ++a) v[a] = __builtin_nontemporal_load(src + (long)a * n4 + i);
        f32x4 s = v[0];
#pragma unroll
        for (int a = 1; a < K; ++a) s = s + v[a];
        __builtin_nontemporal_store(s, dst + i);
    }
}

template <typename TS, typename TD>
__global__ void __launch_bounds__(256) k_cast(const TS* __restrict__ s, TD* __restrict__ d, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) d[i] = (TD)s[i];
}

// phi64 += (double) e over the unknowns (FP64 refinement correction)
template <typename TE>
__global__ void __launch_bounds__(BX * BY)
k_correct(double* __restrict__ phi, const TE* __restrict__ e, Geo g, int dim,
          const int* __restrict__ act)
{
    if (dim == 2 && act != nullptr && act[blockIdx.z] == 0) return;
    const int z = (dim == 3) ? 2 + (int)blockIdx.z : (int)blockIdx.z;
    const int y = 2 + blockIdx.y * BY + threadIdx.y;
    const int x = 2 + blockIdx.x * BX + threadIdx.x;
    if (y > g.my || x > g.mx) return;
    const long p = (long)z * g.sz + (long)y * g.sy + x;
    phi[p] = phi[p] + (double)e[p];
}

// zero the first/last rim+1 layers of the chosen axes (do_BCs with zero
// values, boundary_handler_2d.py:19-29); vals given per face where non-null.
// Axis order as the reference applies it: x faces, then y, then z (3D).
__global__ void __launch_bounds__(256)
k_bcs(double* __restrict__ a, Geo g, int dim, int rim, int direction,
      const double* __restrict__ vxb, const double* __restrict__ vxe,
      const double* __restrict__ vyb, const double* __restrict__ vye, int phase)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (int)(i / ((long)g.nx * g.ny));
    const int w = rim + 1;
    if (dim == 2) {
        if (phase == 0 && (direction == -1 || direction == 0)) {
            if (x < w) a[i] = vxb ? vxb[y] : 0.0;
            else if (x >= g.nx - w) a[i] = vxe ? vxe[y] : 0.0;
        }
        if (phase == 1 && (direction == -1 || direction == 1)) {
            if (y < w) a[i] = vyb ? vyb[x] : 0.0;
            else if (y >= g.ny - w) a[i] = vye ? vye[x] : 0.0;
        }
    } else {
        if (phase == 0 && (direction == -1 || direction == 0)) {
            if (x < w) a[i] = vxb ? vxb[(long)z * g.ny + y] : 0.0;
            else if (x >= g.nx - w) a[i] = vxe ? vxe[(long)z * g.ny + y] : 0.0;
        }
        if (phase == 1 && (direction == -1 || direction == 1)) {
            if (y < w) a[i] = vyb ? vyb[(long)z * g.nx + x] : 0.0;
            else if (y >= g.ny - w) a[i] = vye ? vye[(long)z * g.nx + x] : 0.0;
        }
        if (phase == 2 && (direction == -1 || direction == 2)) {
            // the reference's z faces read vals_yb / vals_ye, indexed [y, x]
            if (z < w) a[i] = vyb ? vyb[(long)y * g.nx + x] : 0.0;
            else if (z >= g.nz - w) a[i] = vye ? vye[(long)y * g.nx + x] : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// Coefficient setup (compute_discretization and helpers).
// ---------------------------------------------------------------------------
struct Src {              // a fine-level input array sampled at level l
    const double* a;      // null -> use scalar
    long sy0, sz0;        // fine strides
    int stride;           // 2^l
    int lvl0;             // level 0 reads the array as is (ghosts included)
};

__device__ __forceinline__ double sample(const Src& s, const Geo& g, int z, int y, int x, int dim)
{
    if (s.lvl0) return s.a[(long)z * s.sz0 + (long)y * s.sy0 + x];
    // injection chain of the reference = strided sample of the fine array on
    // the [1:-1] window, zero on ghosts (solver_2d.py:353, 381)
    if (y < 1 || y > g.my + 1 || x < 1 || x > g.mx + 1) return 0.0;
    if (dim == 3 && (z < 1 || z > g.mz + 1)) return 0.0;
    const long zz = (dim == 3) ? (long)(z - 1) * s.stride + 1 : 0;
    return s.a[zz * s.sz0 + ((long)(y - 1) * s.stride + 1) * s.sy0 + (long)(x - 1) * s.stride + 1];
}

// 2D (solver_2d.py:341-361, 395-421): raw injection at every level, then face
// averages at every level including level 0.
__global__ void __launch_bounds__(256)
k_setup2d(double* __restrict__ cx, double* __restrict__ cy, double* __restrict__ cc,
          double* __restrict__ drr, double* __restrict__ dlr, Geo g,
          Src Dx, Src Dy, Src Lam, Src Dxy, double sqx, double sqy, double cxy,
          double lam_s, double dxs, double dys, int spatial)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int bz = (int)(i / g.sz);            // problem of the batch
    Dx.a = Dx.a ? Dx.a + (long)bz * Dx.sz0 : nullptr;
    Dy.a = Dy.a ? Dy.a + (long)bz * Dy.sz0 : nullptr;
    Lam.a = Lam.a ? Lam.a + (long)bz * Lam.sz0 : nullptr;
    Dxy.a = Dxy.a ? Dxy.a + (long)bz * Dxy.sz0 : nullptr;
    const bool win = (y >= 1 && y <= g.my + 1 && x >= 1 && x <= g.mx + 1);
    const double lam = Lam.a ? (win ? sample(Lam, g, 0, y, x, 2) : 0.0) : lam_s;
    if (!spatial) {
        const double ax = dxs * sqx, ay = dys * sqy;
        cx[i] = ax;
        cy[i] = ay;
        cc[i] = win ? 2.0 * (ax + ay) + lam : 0.0;
        return;
    }
    auto avgx = [&](int yy, int xx) -> double {
        if (yy >= 1 && yy <= g.my + 1 && xx >= 0 && xx <= g.mx + 1)
            return 0.5 * (sample(Dx, g, 0, yy, xx + 1, 2) + sample(Dx, g, 0, yy, xx, 2));
        return sample(Dx, g, 0, yy, xx, 2);
    };
    auto avgy = [&](int yy, int xx) -> double {
        if (yy >= 0 && yy <= g.my + 1 && xx >= 1 && xx <= g.mx + 1)
            return 0.5 * (sample(Dy, g, 0, yy + 1, xx, 2) + sample(Dy, g, 0, yy, xx, 2));
        return sample(Dy, g, 0, yy, xx, 2);
    };
    const double rc = avgx(y, x) * sqx;
    const double cr = avgy(y, x) * sqy;
    cx[i] = rc;
    cy[i] = cr;
    if (win) {
        const double lc = avgx(y, x - 1) * sqx;
        const double cl = avgy(y - 1, x) * sqy;
        cc[i] = lc + rc + cl + cr + lam;
    } else {
        cc[i] = 0.0;
    }
    if (drr != nullptr) {
        if (win) {
            const double q_xp = sample(Dxy, g, 0, y, x + 1, 2);
            const double q_yp = sample(Dxy, g, 0, y + 1, x, 2);
            const double q_xm = sample(Dxy, g, 0, y, x - 1, 2);
            drr[i] = (q_xp + q_yp) * cxy;
            dlr[i] = -(q_xm + q_yp) * cxy;
        } else {
            drr[i] = 0.0;
            dlr[i] = 0.0;
        }
    }
}

// 3D level l >= 1 face-averaged D (solver_3d.py:363-370): injection from the
// previous level's (already averaged, except at level 0) array, then the
// average along `axis` over the reference's window.
__global__ void __launch_bounds__(256)
k_avg3(double* __restrict__ dst, const double* __restrict__ src, Geo g, Geo gs, int axis)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (int)(i / ((long)g.nx * g.ny));
    auto inj = [&](int zz, int yy, int xx) -> double {
        if (zz < 1 || zz > g.mz + 1 || yy < 1 || yy > g.my + 1 || xx < 1 || xx > g.mx + 1) return 0.0;
        return src[(long)(2 * zz - 1) * gs.sz + (long)(2 * yy - 1) * gs.sy + (2 * xx - 1)];
    };
    bool in;
    int dz = 0, dy = 0, dx = 0;
    if (axis == 0) {        // x
        in = (z >= 1 && z <= g.mz + 1 && y >= 1 && y <= g.my + 1 && x <= g.mx + 1);
        dx = 1;
    } else if (axis == 1) { // y
        in = (z >= 1 && z <= g.mz + 1 && y <= g.my + 1 && x >= 1 && x <= g.mx + 1);
        dy = 1;
    } else {                // z
        in = (z <= g.mz + 1 && y >= 1 && y <= g.my + 1 && x >= 1 && x <= g.mx + 1);
        dz = 1;
    }
    dst[i] = in ? 0.5 * (inj(z + dz, y + dy, x + dx) + inj(z, y, x)) : inj(z, y, x);
}

// 3D coefficients of one level from its face arrays (level 0: the raw input,
// never averaged; solver_3d.py:404-429)
__global__ void __launch_bounds__(256)
k_setup3d(double* __restrict__ cx, double* __restrict__ cy, double* __restrict__ cz,
          double* __restrict__ cc, Geo g, const double* __restrict__ ax,
          const double* __restrict__ ay, const double* __restrict__ az, Src Lam,
          double sqx, double sqy, double sqz, double lam_s, double dxs, double dys, double dzs,
          int spatial)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = (int)(i / ((long)g.nx * g.ny));
    const bool win = (z >= 1 && z <= g.mz + 1 && y >= 1 && y <= g.my + 1 && x >= 1 && x <= g.mx + 1);
    const double lam = Lam.a ? (win ? sample(Lam, g, z, y, x, 3) : 0.0) : lam_s;
    if (!spatial) {
        const double axs = dxs * sqx, ays = dys * sqy, azs = dzs * sqz;
        cx[i] = axs;
        cy[i] = ays;
        cz[i] = azs;
        cc[i] = win ? 2.0 * (axs + ays + azs) + lam : 0.0;
        return;
    }
    const double rcc = ax[i] * sqx, crc = ay[i] * sqy, ccr = az[i] * sqz;
    cx[i] = rcc;
    cy[i] = crc;
    cz[i] = ccr;
    if (win) {
        const double lcc = ax[i - 1] * sqx;
        const double clc = ay[i - g.sy] * sqy;
        const double ccl = az[i - g.sz] * sqz;
        cc[i] = lcc + rcc + clc + crc + ccl + ccr + lam;
    } else {
        cc[i] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Non-zero Dirichlet values (the reference's do_BCs with vals_*,
// boundary_handler_2d.py:3-29, which its solver always calls with zeros,
// README.md:90): the values pb holds on the boundary nodes (index 1 and m+1
// per axis) are moved to the right-hand side of the level-0 unknowns next to
// them, rhs_out = rhs - sum over boundary-node neighbours q of A(p, q) pb[q],
// so the zero-boundary solve of rhs_out gives the solution with those values
// (the smoother and the residual see the same sums).  Order per point: x-,
// x+, y-, y+, z-, z+, then the D_xy corners (y-1,x-1), (y-1,x+1), (y+1,x-1),
// (y+1,x+1).  Other entries copy rhs.  2D: z indexes the problems of a batch.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_dirichlet_lift(double* __restrict__ out, const double* __restrict__ rhs, const double* __restrict__ pb,
                 const double* __restrict__ cx, const double* __restrict__ cy, const double* __restrict__ cz,
                 const double* __restrict__ drr, const double* __restrict__ dlr, Geo g, int dim)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const int x = (int)(i % g.nx);
    const int y = (int)((i / g.nx) % g.ny);
    const int z = dim == 3 ? (int)(i / g.sz) : 0;
    double v = rhs[i];
    const bool unk = x >= 2 && x <= g.mx && y >= 2 && y <= g.my && (dim == 2 || (z >= 2 && z <= g.mz));
    if (unk) {
        const long sy = g.sy, sz = g.sz;
        auto bx = [&](int xx) { return xx == 1 || xx == g.mx + 1; };
        auto by = [&](int yy) { return yy == 1 || yy == g.my + 1; };
        if (bx(x - 1)) v = v - cx[i - 1] * pb[i - 1];
        if (bx(x + 1)) v = v - cx[i] * pb[i + 1];
        if (by(y - 1)) v = v - cy[i - sy] * pb[i - sy];
        if (by(y + 1)) v = v - cy[i] * pb[i + sy];
        if (dim == 3) {
            if (z - 1 == 1) v = v - cz[i - sz] * pb[i - sz];
            if (z + 1 == g.mz + 1) v = v - cz[i] * pb[i + sz];
        } else if (drr != nullptr) {
            if (by(y - 1) || bx(x - 1)) v = v - drr[i - sy - 1] * pb[i - sy - 1];
            if (by(y - 1) || bx(x + 1)) v = v - dlr[i - sy + 1] * pb[i - sy + 1];
            if (by(y + 1) || bx(x - 1)) v = v - dlr[i] * pb[i + sy - 1];
            if (by(y + 1) || bx(x + 1)) v = v - drr[i] * pb[i + sy + 1];
        }
    }
    out[i] = v;
}

}  // namespace jsv
