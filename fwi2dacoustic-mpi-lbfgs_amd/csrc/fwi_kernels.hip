// fwi_kernels.hip -- wave-equation core for gfx950 (MI355X).
//
// Three production kernels, one launch per time step for every shot of the
// batch (grid = tiles x shot groups):
//   k_fwd_step   a_lap_opt_2d_pml + time_integral + source + receivers +
//                frame save, fused (fwi_2d.c:761-880, 882-910)
//   k_adj_step   exact transpose of the forward step over the time history,
//                with the residual injected at the receivers
//   k_rec_step   source-field reconstruction on the physical grid
//                (a_lap_opt_2d + time_integral + read_bdr, fwi_2d.c:722-759,
//                866-880, 940-957) fused with the zero-lag imaging condition,
//                summed over the shots of a workgroup in FP64
//
// Tiling: a wave owns 64 adjacent columns (one lane per column) and marches
// down `tz` rows, keeping p (or mu), v and the x-derivative result of the
// last rows in registers.  The x-derivatives of a row go through LDS with a
// 3-cell halo each side; the z-derivatives come from the register window.
// Each workgroup holds up to 4 waves = 4 shots on the same tile.
//
// Arithmetic follows the reference expression by expression (built with
// -ffp-contract=off): float stencils, the free-surface rows and the leapfrog
// sum evaluated in double as C evaluates the reference's double literals.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "fwi_dev.h"

// This file is compiled twice: with the default 8-row strips, and as the
// 4-row variant (FWI_TZ2 = 16, FWI_REC_ROWS = 4, FWI_VARIANT = _s4,
// FWI_SECONDARY).  Each build keeps its kernels in a namespace of its own and
// suffixes the launchers of the strip-height-dependent tile kernels; the
// other launchers come from the default build only.
#ifndef FWI_VARIANT
#define FWI_VARIANT
#endif
#define FWI_CAT2(a, b) a##b
#define FWI_CAT(a, b) FWI_CAT2(a, b)
#define FWI_VNAME(name) FWI_CAT(name, FWI_VARIANT)
#define FWI_KNS FWI_CAT(fwik, FWI_VARIANT)

namespace FWI_KNS {

#define TX FWI_TX

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// CPML table rows: 0 a0x, 1 b0x, 2 a1x, 3 b1x, 4 a1z, 5 b1z; T = 2 npml + 1
#define TB(row, i) tab[(row) * T + (i)]

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2.  The grid is padded to a multiple of 8 and XCD x gets the
// contiguous logical range [x*per, (x+1)*per): neighbouring tiles of a shot,
// which read each other's halo rows, then share an L2.  -1 = padding block.
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int per = (n + 7) >> 3;
    const int l = (b & 7) * per + (b >> 3);
    return l < n ? l : -1;
}

// Raw buffer access: a 128-bit resource (base, size in bytes) in SGPRs, a
// per-lane byte offset (the column, fixed for a wave) and a uniform byte
// offset (the row), so a row costs one scalar add instead of a 64-bit vector
// address.  The resource's range check covers every access: a load whose
// lane offset lies at or beyond the size returns 0 without touching memory
// (buf_ld_row uses that for rows outside a field).
typedef __amdgpu_buffer_rsrc_t brsrc;
#define FWI_BUF_WORD3 0x00020000  // raw 32-bit dwords, gfx9-family resource format
__device__ __forceinline__ brsrc buf_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, FWI_BUF_WORD3);
}
__device__ __forceinline__ float buf_ld(brsrc r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(float v, brsrc r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
// Row r of a field (rows [0, n), pitch rowb bytes) for the general strips, 0
// outside the field.  The load is always issued (a load inside a branch of
// its own is waited for before the branch ends, one memory latency per row);
// for a row outside the field its lane offset is FWI_BUF_OUT, beyond any
// resource size, so the range check returns 0.  Selecting 0 after the load
// instead keeps one 64-bit lane mask per row alive until the loads return:
// with every load of a strip issued first, that spilled 70 to 200 SGPRs of
// the general bodies into VGPR lanes.
// (n >= 0: one unsigned compare tests 0 <= r < n)
__device__ __forceinline__ float buf_ld_row(brsrc rs, int voff, int r, int n, int rowb, int base = 0)
{
    const bool in = (unsigned)r < (unsigned)n;
    return buf_ld(rs, in ? voff : FWI_BUF_OUT, base + (in ? r : 0) * rowb);
}
// the same with 0 selected after the load: one VGPR for all rows' lane offset
// (the reconstruction items, whose loads of two shots are in flight at once,
// run out of VGPRs with an offset per row)
__device__ __forceinline__ float buf_ld_row_sel(brsrc rs, int voff, int r, int n, int rowb)
{
    const bool in = (unsigned)r < (unsigned)n;
    const float v = buf_ld(rs, voff, (in ? r : 0) * rowb);
    return in ? v : 0.f;
}
// A pointer read from the parameter block is generic to the compiler, so its
// loads and stores become flat instructions, which count against the LDS
// wait counter as well: a wait for the stencil's LDS rows then also waits for
// every such load in flight.  Every buffer of the context is device memory,
// so such pointers are taken into the global address space, where the
// compiler emits global loads and stores.
#define FWI_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ FWI_GLOBAL T *gbl(T *p)
{
    return (FWI_GLOBAL T *)p;
}
// element i of a, 0 when !in; the load is issued for element 0 then (a must
// hold at least one element), for the reason above
template <class T>
__device__ __forceinline__ T ld_sel(const FWI_GLOBAL T *a, long long i, bool in, T zero)
{
    const T v = a[in ? i : 0];
    return in ? v : zero;
}
template <class T>
__device__ __forceinline__ T ld_sel(const T *a, long long i, bool in, T zero)
{
    return ld_sel(gbl(a), i, in, zero);
}

// A uniform value held in a scalar register from here on.  Where a lane
// condition picks between two fields of the parameter block (which lives in
// device memory), the compiler would otherwise read the picked field per lane
// with a vector load, and wait for it inside the divergent branch.
__device__ __forceinline__ float uni(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
// the x-stencil constants of the propagation grid, in registers
struct XConst {
    int pnx, npml;
    float ix, c0x, c1x;
};
__device__ __forceinline__ XConst xconst(const FwiGeomDev &g)
{
    return XConst{uni(g.pnx), uni(g.npml), uni(g.ix), uni(g.c0x), uni(g.c1x)};
}

// frame position of physical cell (bz, bx) inside one step of the strips
__device__ __forceinline__ int frame_pos(int bz, int bx, int nx, int nz)
{
    if (bz < 3) return bz * nx + bx;
    if (bz >= nz - 3) return (3 + nz - 1 - bz) * nx + bx;
    if (bx < 3) return 6 * nx + bz * 6 + bx;
    return 6 * nx + bz * 6 + 3 + (nx - 1 - bx);
}

// ============================================================================
// forward: x chain of one row (u nodes, CPML on u, ux nodes, CPML on ux)
// ============================================================================

// Column that lane `lane` (< 6) loads for the x halo of a 64-column tile:
// lanes 0..2 -> x0-3..x0-1, lanes 3..5 -> x0+64..x0+66 (-1: no halo lane).
__device__ __forceinline__ int halo_col(int x0, int lane)
{
    return (lane < 3) ? x0 - 3 + lane : ((lane < 6) ? x0 + 61 + lane : -1);
}

// Rows per load block of the marching kernels: a wave issues the loads of
// RB rows (own column, halo, in-place field, ekp) before computing any of
// them, so RB rows of memory traffic are in flight per wave instead of one.
#ifndef FWI_ROWBLOCK
#define FWI_ROWBLOCK 4
#endif
#define RB FWI_ROWBLOCK

// 1: the main blocks of a march load block b+1 while block b is computed
#ifndef FWI_DOUBLEBUF
#define FWI_DOUBLEBUF 1
#endif

// CPML tables staged in LDS once per workgroup (6 rows of 2 npml + 1)
#define FWI_MAXNPML 64
#define FWI_TABN (6 * (2 * FWI_MAXNPML + 1))

// any workgroup size (the z-marching kernels)
__device__ __forceinline__ void stage_tables(float *stab, const float *tab, int npml)
{
    const int n = 6 * (2 * npml + 1);
    for (int i = threadIdx.x; i < n; i += blockDim.x) stab[i] = gbl(tab)[i];
    __syncthreads();
}
// 256-thread workgroups (the tile kernels): every load issued before the first
// LDS store, so the staging costs one memory latency instead of one per pass
__device__ __forceinline__ void stage_tables_256(float *stab, const float *tab, int npml)
{
    constexpr int NP = (FWI_TABN + 255) / 256;
    const int n = 6 * (2 * npml + 1);
    float t[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int i = threadIdx.x + 256 * q;
        t[q] = gbl(tab)[i < n ? i : 0];
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int i = threadIdx.x + 256 * q;
        if (i < n) stab[i] = t[q];
    }
    __syncthreads();
}

// u at half node j (between p[j-1] and p[j]); pi = LDS index of p[j]
__device__ __forceinline__ float fwd_u_node(const FwiGeomDev &g, const float *tab, const float *Ps,
                                            int j, int pi, int r, const float *xh_o, float *xh_n,
                                            int x0)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1, n1 = npml + 1;
    float u;
    if (j <= 0 || j >= pnx) u = 0.f;
    else if (j == 1 || j == pnx - 1) u = g.ix * (Ps[pi] - Ps[pi - 1]);
    else u = g.c0x * (Ps[pi] - Ps[pi - 1]) + g.c1x * (Ps[pi + 1] - Ps[pi - 2]);
    const bool own = (j >= x0) && (j < x0 + TX);
    if (j >= 0 && j <= npml) {
        float ps = xh_o[r * 2 * n1 + j];
        ps = TB(1, 2 * j) * ps + TB(0, 2 * j) * u;
        u = u + ps;
        if (own) xh_n[r * 2 * n1 + j] = ps;
    }
    if (j >= pnx - npml && j <= pnx) {
        const int q = j - (pnx - npml);
        float ps = xh_o[r * 2 * n1 + n1 + q];
        ps = TB(3, 2 * q) * ps + TB(2, 2 * q) * u;
        u = u + ps;
        if (own) xh_n[r * 2 * n1 + n1 + q] = ps;
    }
    return u;
}

// ux at p node i; ui = LDS index of u[i]
__device__ __forceinline__ float fwd_ux_node(const FwiGeomDev &g, const float *tab, const float *Us,
                                             int i, int ui, int r, const float *xd_o, float *xd_n)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    if (i >= pnx) return 0.f;
    float ux;
    if (i == 0 || i == pnx - 1) ux = (Us[ui + 1] - Us[ui]) * g.ix;
    else ux = g.c0x * (Us[ui + 1] - Us[ui]) + g.c1x * (Us[ui + 2] - Us[ui - 1]);
    if (i < npml) {
        float ps = xd_o[r * 2 * npml + i];
        ps = TB(1, 2 * i + 1) * ps + TB(0, 2 * i + 1) * ux;
        ux = ux + ps;
        xd_n[r * 2 * npml + i] = ps;
    } else if (i >= pnx - npml) {
        const int q = i - (pnx - npml);
        float ps = xd_o[r * 2 * npml + npml + q];
        ps = TB(3, 2 * q + 1) * ps + TB(2, 2 * q + 1) * ux;
        ux = ux + ps;
        xd_n[r * 2 * npml + npml + q] = ps;
    }
    return ux;
}

// x chain of a block of RB rows: all rows staged in LDS, then all u nodes,
// then all ux nodes (3 wave syncs per block).  Rows outside [z0, z1) are
// skipped (their ux is 0 and never used).
template <bool FAST, int N>
__device__ __forceinline__ void fwd_xblock(const FwiGeomDev &g, const float *tab, const float *bs, const float *bh,
                                           float (*Ps)[72], float (*Us)[72], int x0, int lane, int rb, int z0,
                                           int z1, const float *xh_o, float *xh_n, const float *xd_o, float *xd_n,
                                           float *bux)
{
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        if (r >= z0 && r < z1) {
            Ps[d][lane + 3] = bs[d];
            if (lane < 3) Ps[d][lane] = bh[d];
            else if (lane < 6) Ps[d][lane + 64] = bh[d];
        }
    }
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        if (r >= z0 && r < z1) {
            const float *P = Ps[d];
            float *U = Us[d];
            if (FAST) {
                U[lane] = g.c0x * (P[lane + 2] - P[lane + 1]) + g.c1x * (P[lane + 3] - P[lane]);
                if (lane < 3)
                    U[64 + lane] = g.c0x * (P[lane + 66] - P[lane + 65]) + g.c1x * (P[lane + 67] - P[lane + 64]);
            } else {
                U[lane] = fwd_u_node(g, tab, P, x0 - 1 + lane, lane + 2, r, xh_o, xh_n, x0);
                if (lane < 3) U[64 + lane] = fwd_u_node(g, tab, P, x0 + 63 + lane, lane + 66, r, xh_o, xh_n, x0);
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        bux[d] = 0.f;
        if (r >= z0 && r < z1) {
            const float *U = Us[d];
            bux[d] = FAST ? g.c0x * (U[lane + 2] - U[lane + 1]) + g.c1x * (U[lane + 3] - U[lane])
                          : fwd_ux_node(g, tab, U, x0 + lane, lane + 1, r, xd_o, xd_n);
        }
    }
}

__device__ __forceinline__ float v_row0(float p0, float p1, float p2, float p3, float iz840)
{
    return (float)((3675.0 * p0 - 1225.0 * p1 + 441.0 * p2 - 75.0 * p3) * (double)iz840);
}
__device__ __forceinline__ float v_row1(float p0, float p1, float p2, float p3, float iz840)
{
    return (float)((-1085.0 * p0 + 1015.0 * p1 - 63.0 * p2 + 5.0 * p3) * (double)iz840);
}
__device__ __forceinline__ float vz_row0(float v0, float v1, float v2, float v3, float v4, float iz24)
{
    return (float)((-22. * v0 + 17. * v1 + 9. * v2 - 5. * v3 + v4) * (double)iz24);
}

// leapfrog, fwi_2d.c:878: 2.0*pnow - pold + ekp*(pxx+pzz), in double, rounded once
__device__ __forceinline__ float leapfrog(float pnow, float pold, float k_times_x)
{
    return (float)(2.0 * (double)pnow - (double)pold + (double)k_times_x);
}

__global__ __launch_bounds__(FWI_TX * FWI_MAXSG) void k_fwd_step(FwiFwdArgs A, const int *kbase, int koff)
{
    // step index: koff alone, or a device counter plus koff inside a captured graph
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiGeomDev &g = A.g;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ntiles = g.ntx * g.ntz;
    const int lb = xcd_remap(blockIdx.x, ntiles * A.b.ngroups);
    if (lb < 0) return;
    const int grp = lb / ntiles, t = lb - grp * ntiles;
    const int tzi = t / g.ntx, txi = t - tzi * g.ntx;
    const int s = grp * A.b.sg + wv;
    __shared__ float stab[FWI_TABN];
    stage_tables(stab, A.b.tab, g.npml);
    if (s >= A.b.nshots) return;
    const FwiShotDev sh = A.b.shots[s];
    if (k >= sh.nt) return;

    __shared__ float lds[FWI_MAXSG][2][RB][72];
    float(*Ps)[72] = lds[wv][0];
    float(*Us)[72] = lds[wv][1];

    const int npml = g.npml, pnx = g.pnx, pnz = g.pnz, ld = g.ld, T = 2 * npml + 1;
    const float *tab = stab;
    const float fsrc = A.b.srcf[sh.src_off + k];
    const int x0 = txi * TX, x = x0 + lane;
    const int z0 = tzi * g.tz, z1 = min(z0 + g.tz, pnz);
    const long long so = (long long)s * A.b.fs;
    const int cur = k & 1, nxt = cur ^ 1;
    const float *pnow = (cur ? A.P1 : A.P0) + so;
    float *pold = (cur ? A.P0 : A.P1) + so;
    const float *ekp = A.b.ekp;
    const float *xh_o = A.xh + s * A.b.xh_s + cur * A.b.xh_b;
    float *xh_n = A.xh + s * A.b.xh_s + nxt * A.b.xh_b;
    const float *xd_o = A.xd + s * A.b.xd_s + cur * A.b.xd_b;
    float *xd_n = A.xd + s * A.b.xd_s + nxt * A.b.xd_b;
    const float *zh_o = A.zh + s * A.b.zh_s + cur * A.b.zh_b;
    float *zh_n = A.zh + s * A.b.zh_s + nxt * A.b.zh_b;
    const float *zd_o = A.zd + s * A.b.zd_s + cur * A.b.zd_b;
    float *zd_n = A.zd + s * A.b.zd_s + nxt * A.b.zd_b;
    const bool fastx = (x0 >= npml + 3) && (x0 + 66 <= pnx - npml - 1);
    const int bx = x - npml, nx = g.nx, nz = g.nz;
    const bool in_bx = (bx >= 0) && (bx < nx);

    float p0w = 0.f, p1w = 0.f, p2w = 0.f, p3w = 0.f;  // p[r-3..r]
    float v0w = 0.f, v1w = 0.f, v2w = 0.f, v3w = 0.f;  // v[r-4..r-1]
    float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;      // ux[r-3..r]

    // iteration r needs p[r] (own column and halo) and, for its output row
    // i = r - 3, p^{k-1}[i] and ekp[i]; loaded a block of RB rows at a time
    const int rend = z1 + 2;
    const int hq = halo_col(x0, lane);
    auto fetch = [&](int r, float &fs, float &fh, float &fo, float &fk) {
        fs = fh = fo = fk = 0.f;
        if (r > rend) return;
        if (r >= 0 && r < pnz) fs = pnow[(long long)r * ld + x];
        if (r >= z0 && r < z1 && hq >= 0 && hq < pnx) fh = pnow[(long long)r * ld + hq];
        const int io = r - 3;
        if (io >= z0 && io < z1) {
            fo = pold[(long long)io * ld + x];
            fk = ekp[(long long)io * ld + x];
        }
    };
    auto body = [&](int r, float pr, float uxr, float pold_i, float ekp_i) {
        p0w = p1w; p1w = p2w; p2w = p3w; p3w = pr;
        q0 = q1; q1 = q2; q2 = q3; q3 = uxr;

        // v at half node j = r - 1: window holds p[j-2..j+1]
        const int j = r - 1;
        float vj = 0.f;
        if (j >= 0 && j < pnz) {
            if (j == 0) vj = v_row0(p2w, p3w, pnow[2 * ld + x], pnow[3 * ld + x], g.iz840);
            else if (j == 1) vj = v_row1(p1w, p2w, p3w, pnow[3 * ld + x], g.iz840);
            else if (j == pnz - 1) vj = (p2w - p1w) * g.iz;
            else vj = g.c0z * (p2w - p1w) + g.c1z * (p3w - p0w);
            if (j >= pnz - npml) {
                const int q = j - (pnz - npml);
                float ps = zh_o[q * ld + x];
                ps = TB(5, 2 * q) * ps + TB(4, 2 * q) * vj;
                vj = vj + ps;
                if (j >= z0 && j < z1) zh_n[q * ld + x] = ps;
            }
        }
        v0w = v1w; v1w = v2w; v2w = v3w; v3w = vj;

        const int i = r - 3;  // output row: window v[i-1..i+2], p[i] = p0w, ux[i] = q0
        if (i >= z0 && i < z1) {
            float vz;
            if (i == 0) {
                const float p4 = pnow[4 * ld + x], p5 = pnow[5 * ld + x];
                const float v3 = g.c0z * (p3w - p2w) + g.c1z * (p4 - p1w);
                const float v4 = g.c0z * (p4 - p3w) + g.c1z * (p5 - p2w);
                vz = vz_row0(v1w, v2w, v3w, v3, v4, g.iz24);
            } else if (i == pnz - 1) {
                vz = (v2w - v1w) * g.iz;
            } else {
                vz = g.c0z * (v2w - v1w) + g.c1z * (v3w - v0w);
            }
            if (i >= pnz - npml) {
                const int q = i - (pnz - npml);
                float ps = zd_o[q * ld + x];
                ps = TB(5, 2 * q + 1) * ps + TB(4, 2 * q + 1) * vz;
                vz = vz + ps;
                zd_n[q * ld + x] = ps;
            }
            float ux = q0;
            if (i == sh.isz && x == sh.isx) ux = ux + fsrc;
            const float X = ux + vz;
            const long long o = (long long)i * ld + x;
            const float pnew = leapfrog(p0w, pold_i, ekp_i * X);
            if (x < pnx) pold[o] = pnew;
            if (i >= sh.rz_lo && i <= sh.rz_hi) {
                const int slot = A.b.rrow[sh.rrow_off + i];
                if (slot >= 0) {
                    const int ri = gbl(A.b.rmap)[slot * ld + x];
                    if (ri >= 0) A.syn[sh.trace_off + (long long)k * sh.nr + ri] = pnew;
                }
            }
            if (in_bx && i < nz && (i < 3 || i >= nz - 3 || bx < 3 || bx >= nx - 3)) {
                const long long so2 = sh.strip_off + (long long)k * g.sl;
                if (i < 3 || i >= nz - 3) {
                    const int pos = (i < 3) ? i * nx + bx : (3 + nz - 1 - i) * nx + bx;
                    A.pstrip[so2 + pos] = pnew;
                    A.xstrip[so2 + pos] = X;
                }
                if (bx < 3 || bx >= nx - 3) {
                    const int pos = 6 * nx + i * 6 + ((bx < 3) ? bx : 3 + (nx - 1 - bx));
                    A.pstrip[so2 + pos] = pnew;
                    A.xstrip[so2 + pos] = X;
                }
            }
        }
    };
    // interior rows: no special row, strip, frame, source or receiver
    auto plain_body = [&](int r, float pr, float uxr, float pold_i, float ekp_i) {
        p0w = p1w; p1w = p2w; p2w = p3w; p3w = pr;
        q0 = q1; q1 = q2; q2 = q3; q3 = uxr;
        const float vj = g.c0z * (p2w - p1w) + g.c1z * (p3w - p0w);
        v0w = v1w; v1w = v2w; v2w = v3w; v3w = vj;
        const float vz = g.c0z * (v2w - v1w) + g.c1z * (v3w - v0w);
        const float X = q0 + vz;
        pold[(r - 3) * ld + x] = leapfrog(p0w, pold_i, ekp_i * X);
    };
    // any N rows: general x chain and bodies
    auto gen_block = [&](auto Nc, int rb) {
        constexpr int N = decltype(Nc)::value;
        float bs[N], bh[N], bo[N], bk[N], bu[N];
#pragma unroll
        for (int d = 0; d < N; ++d) fetch(rb + d, bs[d], bh[d], bo[d], bk[d]);
        if (rb + N - 1 >= z0 && rb < z1) {
            if (fastx) fwd_xblock<true, N>(g, tab, bs, bh, Ps, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
            else fwd_xblock<false, N>(g, tab, bs, bh, Ps, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
        } else {
#pragma unroll
            for (int d = 0; d < N; ++d) bu[d] = 0.f;
        }
#pragma unroll
        for (int d = 0; d < N; ++d)
            if (rb + d <= rend) body(rb + d, bs[d], bu[d], bo[d], bk[d]);
    };
    // a main block: RB rows that all need the x chain and all output a row
    auto run_block = [&](int rb, const float *bs, const float *bh, const float *bo, const float *bk) {
        float bu[RB];
        const int ilo = rb - 3, ihi = rb + RB - 4;
        const bool plain = fastx && ilo >= 3 && ihi < nz - 3 && (sh.isz < ilo || sh.isz > ihi) &&
                           (sh.rz_hi < ilo || sh.rz_lo > ihi);
        if (plain) {
            fwd_xblock<true, RB>(g, tab, bs, bh, Ps, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
#pragma unroll
            for (int d = 0; d < RB; ++d) plain_body(rb + d, bs[d], bu[d], bo[d], bk[d]);
        } else {
            if (fastx) fwd_xblock<true, RB>(g, tab, bs, bh, Ps, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
            else fwd_xblock<false, RB>(g, tab, bs, bh, Ps, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
#pragma unroll
            for (int d = 0; d < RB; ++d) body(rb + d, bs[d], bu[d], bo[d], bk[d]);
        }
    };
    using I3 = std::integral_constant<int, 3>;
    // warm-up rows z0-3 .. z0+2 (no output yet)
    gen_block(I3{}, z0 - 3);
    gen_block(I3{}, z0);
    // main rows z0+3 .. z1-1, double-buffered: the loads of the next block
    // are in flight while the current block is computed
    int rb = z0 + 3;
    const int rmain = z1 - 1;
    if (rb + RB - 1 <= rmain) {
        float As[RB], Ah[RB], Ao[RB], Ak[RB], Bs[RB], Bh[RB], Bo[RB], Bk[RB];
#pragma unroll
        for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ao[d], Ak[d]);
        for (;;) {
            const bool nb = rb + 2 * RB - 1 <= rmain;
            if (nb && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, Bs[d], Bh[d], Bo[d], Bk[d]);
            }
            run_block(rb, As, Ah, Ao, Ak);
            rb += RB;
            if (!nb) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, Bs[d], Bh[d], Bo[d], Bk[d]);
            }
            const bool na = rb + 2 * RB - 1 <= rmain;
            if (na && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, As[d], Ah[d], Ao[d], Ak[d]);
            }
            run_block(rb, Bs, Bh, Bo, Bk);
            rb += RB;
            if (!na) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ao[d], Ak[d]);
            }
        }
    }
    // tail rows up to z1+2
    for (; rb <= rend; rb += 3) gen_block(I3{}, rb);
}

// ============================================================================
// forward, 2-D tile form: one workgroup (4 waves) per 64 x FWI_TZ2 tile of one
// shot; wave w owns the FR contiguous output rows r0 = z0 + w FR ...  A wave
// loads the FR + 6 rows of p its rows depend on (3-row halo above and below)
// into registers, so v (z-derivative of p) and vz (z-derivative of v) never
// leave registers; only the x chain (u, then ux) goes through LDS, on rows
// private to the wave (wave-level sync only).  Every load of a wave is issued
// before any arithmetic.  Tiles that touch no CPML strip, free-surface row,
// frame strip, source or receiver row run the interior body, which has no
// lane-dependent branch; the others run the general body.  Same arithmetic
// and CPML ownership rules as k_fwd_step.
// ============================================================================
#define FR (FWI_TZ2 / 4) // output rows per wave
#define FPR (FR + 6)      // p rows per wave: r0-3 .. r0+FR+2
#define FVN (FR + 3)      // v nodes per wave: r0-1 .. r0+FR+1
#define HC 72             // LDS row: 3 halo | 64 | 3 halo | 2 spare
#define HJ 71             // spare column that takes the stores of idle lanes
#define HP ((6 * FR + 63) / 64)  // x-halo loads per wave: 6 values per own row
static_assert(FWI_TZ2 % 4 == 0 && FR >= 4, "a tile is 4 waves of FR rows");
static_assert(3 * FR <= 64, "one pass for the three extra u nodes of every row");

// The x halo of a wave's nrow rows: value h of lane l is number l + 64 h =
// 6 d + c of the list (row d, column c: c < 3 left of the tile, else right).
template <int NP>
struct XHalo {
    int d[NP], col[NP];  // row within the wave, column (grid)
    bool in[NP];         // a real entry (not past the 6 nrow values)
};
template <int NP>
__device__ __forceinline__ void xhalo_plan(XHalo<NP> &h, int lane, int x0, int nrow)
{
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int l = lane + 64 * q, d = l / 6, c = l - 6 * d;
        h.in[q] = d < nrow;
        h.d[q] = h.in[q] ? d : 0;
        h.col[q] = (c < 3) ? x0 - 3 + c : x0 + 61 + c;
    }
}
// LDS column of entry q (the spare column for idle entries)
template <int NP>
__device__ __forceinline__ int xhalo_lds_col(const XHalo<NP> &h, int q, int x0)
{
    const int c = h.col[q] - x0 + 3;
    return h.in[q] ? c : HJ;
}

// CPML memory index of u node j (x strips, row r), or -1 outside the strips
__device__ __forceinline__ int xh_idx(const FwiGeomDev &g, int j, int r)
{
    const int npml = g.npml, n1 = npml + 1;
    if (j >= 0 && j <= npml) return r * 2 * n1 + j;
    if (j >= g.pnx - npml && j <= g.pnx) return r * 2 * n1 + n1 + (j - (g.pnx - npml));
    return -1;
}
// CPML memory index of p node i (x strips, row r), or -1 outside the strips
__device__ __forceinline__ int xd_idx(const FwiGeomDev &g, int i, int r)
{
    const int npml = g.npml;
    if (i >= 0 && i < npml) return r * 2 * npml + i;
    if (i >= g.pnx - npml && i < g.pnx) return r * 2 * npml + npml + (i - (g.pnx - npml));
    return -1;
}
// u at half node j with the CPML update from the old memory value po (new in pn):
// the arithmetic of fwd_u_node, memory traffic left to the caller
__device__ __forceinline__ float fwd_u_edge(const FwiGeomDev &g, const float *tab, const float *Ps, int j, int pi,
                                            float po, float &pn)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    float u;
    if (j <= 0 || j >= pnx) u = 0.f;
    else if (j == 1 || j == pnx - 1) u = g.ix * (Ps[pi] - Ps[pi - 1]);
    else u = g.c0x * (Ps[pi] - Ps[pi - 1]) + g.c1x * (Ps[pi + 1] - Ps[pi - 2]);
    pn = 0.f;
    if (j >= 0 && j <= npml) {
        pn = TB(1, 2 * j) * po + TB(0, 2 * j) * u;
        u = u + pn;
    } else if (j >= pnx - npml && j <= pnx) {
        const int q = j - (pnx - npml);
        pn = TB(3, 2 * q) * po + TB(2, 2 * q) * u;
        u = u + pn;
    }
    return u;
}
// ux at p node i (ui = LDS index of u[i]) with the CPML update: fwd_ux_node's arithmetic
__device__ __forceinline__ float fwd_ux_edge(const FwiGeomDev &g, const float *tab, const float *Us, int i, int ui,
                                             float po, float &pn)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    pn = 0.f;
    if (i >= pnx) return 0.f;
    float ux;
    if (i == 0 || i == pnx - 1) ux = (Us[ui + 1] - Us[ui]) * g.ix;
    else ux = g.c0x * (Us[ui + 1] - Us[ui]) + g.c1x * (Us[ui + 2] - Us[ui - 1]);
    if (i < npml) {
        pn = TB(1, 2 * i + 1) * po + TB(0, 2 * i + 1) * ux;
        ux = ux + pn;
    } else if (i >= pnx - npml) {
        const int q = i - (pnx - npml);
        pn = TB(3, 2 * q + 1) * po + TB(2, 2 * q + 1) * ux;
        ux = ux + pn;
    }
    return ux;
}


// Branch-free x chain of the general tiles: per-node coefficients, computed
// once per wave, that reproduce fwd_u_edge / fwd_ux_edge exactly.  The
// stencil is cA (f[+0] - f[-1]) + cB (f[+1] - f[-2]) (cB = 0 on the 2nd-order
// edge nodes, both 0 outside the grid, so the added terms are exact zeros);
// the CPML update is pn = cb po + ca x, x + pn (ca = cb = 0 outside the strips).
struct XCoef {
    float A, B, ca, cb;
    int col;  // column of the node's CPML memory within its row (-1: none)
};
// u node j (between p[j-1] and p[j])
__device__ __forceinline__ XCoef fwd_u_coef(const XConst &g, const float *tab, int j)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1, n1 = npml + 1;
    XCoef c;
    c.A = (j <= 0 || j >= pnx) ? 0.f : ((j == 1 || j == pnx - 1) ? g.ix : g.c0x);
    c.B = (j <= 1 || j >= pnx - 1) ? 0.f : g.c1x;
    c.ca = c.cb = 0.f;
    c.col = -1;
    if (j >= 0 && j <= npml) {
        c.cb = TB(1, 2 * j);
        c.ca = TB(0, 2 * j);
        c.col = j;
    } else if (j >= pnx - npml && j <= pnx) {
        const int q = j - (pnx - npml);
        c.cb = TB(3, 2 * q);
        c.ca = TB(2, 2 * q);
        c.col = n1 + q;
    }
    return c;
}
// ux node i (p node)
__device__ __forceinline__ XCoef fwd_ux_coef(const XConst &g, const float *tab, int i)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    XCoef c;
    c.A = (i < 0 || i >= pnx) ? 0.f : ((i == 0 || i == pnx - 1) ? g.ix : g.c0x);
    c.B = (i <= 0 || i >= pnx - 1) ? 0.f : g.c1x;
    c.ca = c.cb = 0.f;
    c.col = -1;
    if (i >= 0 && i < npml) {
        c.cb = TB(1, 2 * i + 1);
        c.ca = TB(0, 2 * i + 1);
        c.col = i;
    } else if (i >= pnx - npml && i < pnx) {
        const int q = i - (pnx - npml);
        c.cb = TB(3, 2 * q + 1);
        c.ca = TB(2, 2 * q + 1);
        c.col = npml + q;
    }
    return c;
}

template <bool EDGE, int MODE, int CLS = FWI_CLS_ALL>
__device__ __forceinline__ void fwd_tile_rows(const FwiFwdArgs &A, const FwiShotDev &sh, const float *tab,
                                              float (*Ps)[HC], float (*Us)[HC], int s, int k, int x0, int r0,
                                              int lane)
{
    const FwiGeomDev &g = A.g;
    const int npml = g.npml, pnx = g.pnx, pnz = g.pnz, ld = g.ld, T = 2 * npml + 1;
    const int x = x0 + lane;
    const long long so = (long long)s * A.b.fs;
    const int cur = k & 1, nxt = cur ^ 1;
    const float *pnow = (cur ? A.P1 : A.P0) + so;
    float *pold = (cur ? A.P0 : A.P1) + so;
    const unsigned fbytes = (unsigned)(A.b.fs * sizeof(float));
    const brsrc rnow = buf_rsrc(pnow, fbytes), rold = buf_rsrc(pold, fbytes), rk = buf_rsrc(A.b.ekp, fbytes);
    const int vo = x * 4, rowb = ld * 4;  // byte offsets: column (per lane), row pitch

    // ---- every load of the wave
    float p[FPR], po[FR], pk[FR];
#pragma unroll
    for (int m = 0; m < FPR; ++m) {
        const int zr = r0 - 3 + m;
        p[m] = EDGE ? buf_ld_row(rnow, vo, zr, pnz, rowb) : buf_ld(rnow, vo, zr * rowb);
    }
    // x halo of the FR own rows, HP loads
    XHalo<HP> xh;
    xhalo_plan(xh, lane, x0, FR);
    float hv[HP];
#pragma unroll
    for (int q = 0; q < HP; ++q) {
        bool ok = xh.in[q];
        if (EDGE) ok = ok && r0 + xh.d[q] < pnz && xh.col[q] >= 0 && xh.col[q] < pnx;
        hv[q] = buf_ld(rnow, ok ? (xh.d[q] * ld + xh.col[q]) * 4 : vo, r0 * rowb);
        hv[q] = ok ? hv[q] : 0.f;
    }
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        const int i = r0 + d;
        po[d] = EDGE ? buf_ld_row(rold, vo, i, pnz, rowb) : buf_ld(rold, vo, i * rowb);
        pk[d] = EDGE ? buf_ld_row(rk, vo, i, pnz, rowb) : buf_ld(rk, vo, i * rowb);
    }

    const FWI_GLOBAL float *xh_o = gbl(A.xh) + s * A.b.xh_s + cur * A.b.xh_b;
    FWI_GLOBAL float *xh_n = gbl(A.xh) + s * A.b.xh_s + nxt * A.b.xh_b;
    const FWI_GLOBAL float *xd_o = gbl(A.xd) + s * A.b.xd_s + cur * A.b.xd_b;
    FWI_GLOBAL float *xd_n = gbl(A.xd) + s * A.b.xd_s + nxt * A.b.xd_b;
    const FWI_GLOBAL float *zh_o = gbl(A.zh) + s * A.b.zh_s + cur * A.b.zh_b;
    FWI_GLOBAL float *zh_n = gbl(A.zh) + s * A.b.zh_s + nxt * A.b.zh_b;
    const FWI_GLOBAL float *zd_o = gbl(A.zd) + s * A.b.zd_s + cur * A.b.zd_b;
    FWI_GLOBAL float *zd_n = gbl(A.zd) + s * A.b.zd_s + nxt * A.b.zd_b;
    // u nodes x0-1 .. x0+65 and p nodes x0 .. x0+63 clear of the x CPML and edges
    // (each special rule is compiled in only when the body's mask CLS allows it)
    const bool fastx = !EDGE || !(CLS & FWI_CLS_X) || ((x0 >= npml + 3) && (x0 + 66 <= pnx - npml - 1));
    const int ud0 = lane / 3;  // the extra-u-node pass: lane = 3 ud + (uc - 64)
    bool uok = lane < 3 * FR;
    if (EDGE) uok = uok && r0 + ud0 < pnz;
    // idle lanes evaluate node x0+64 of row 0 (owned by the next tile, so no CPML store)
    const int ud = uok ? ud0 : 0, uc = uok ? 64 + lane - 3 * ud0 : 65;

    // ---- general tiles: every other load before the first store.  Vector
    // memory operations retire in order, so a load issued after a store waits
    // for that store; the CPML memory, receiver map and source sample are
    // therefore read here, not where they are used.
    float zho[FVN], zdo[FR], xuo[FR], xuho = 0.f, xdo[FR], sv = 0.f;
    int rix[FR];
    XCoef cu, cuh, cx;  // x chain coefficients of this lane's u node, extra u node and ux node
    const int n1 = npml + 1;
    // which special rules can occur in this strip at all (uniform): each is
    // tested once here instead of per row and node
    const bool zplain = !EDGE || !(CLS & FWI_CLS_Z) || (r0 >= 3 && r0 + FR + 1 < pnz - npml);  // plain v, vz
    constexpr bool TRACES = MODE == FWI_FWD_FRAMES || MODE == FWI_FWD_TRACES;  // receivers recorded
    const bool has_recv =
        TRACES && EDGE && (CLS & FWI_CLS_RS) && sh.rz_lo <= sh.rz_hi && sh.rz_lo < r0 + FR && sh.rz_hi >= r0;
    const bool has_src =
        EDGE && (CLS & FWI_CLS_RS) && sh.isz >= r0 && sh.isz < r0 + FR && sh.isx >= x0 && sh.isx < x0 + TX;
    const bool frame_rows = EDGE && (CLS & FWI_CLS_FR) && (r0 < 3 || (r0 + FR > g.nz - 3 && r0 < g.nz));
    const bool frame_cols = EDGE && (CLS & FWI_CLS_FR) &&
                            ((x0 < npml + 3 && x0 + TX > npml) || (x0 < npml + g.nx && x0 + TX > npml + g.nx - 3));
    if (EDGE && !fastx) {
        const XConst xk = xconst(g);
        cu = fwd_u_coef(xk, tab, x0 - 1 + lane);
        cuh = fwd_u_coef(xk, tab, x0 - 1 + uc);
        cx = fwd_ux_coef(xk, tab, x);
    }
    if (EDGE) {
        if (!zplain) {
#pragma unroll
            for (int n = 0; n < FVN; ++n) {
                const int j = r0 - 1 + n;
                zho[n] = ld_sel(zh_o, (long long)(j - (pnz - npml)) * ld + x, j >= pnz - npml && j < pnz, 0.f);
            }
#pragma unroll
            for (int d = 0; d < FR; ++d) {
                const int i = r0 + d;
                zdo[d] = ld_sel(zd_o, (long long)(i - (pnz - npml)) * ld + x, i >= pnz - npml && i < pnz, 0.f);
            }
        }
#pragma unroll
        for (int d = 0; d < FR; ++d) {
            // x strips: per lane, clamped index, no divergence; rows past the
            // grid read the last row (their nodes are not evaluated)
            const int i = min(r0 + d, pnz - 1);
            xuo[d] = xdo[d] = 0.f;
            if (!fastx) {
                xuo[d] = xh_o[i * 2 * n1 + (cu.col >= 0 ? cu.col : 0)];
                xdo[d] = xd_o[i * 2 * npml + (cx.col >= 0 ? cx.col : 0)];
            }
        }
#pragma unroll
        for (int d = 0; d < FR; ++d) rix[d] = -1;
        if (has_recv) {  // every row slot first, then the lanes' receiver indices
            int slot[FR];
#pragma unroll
            for (int d = 0; d < FR; ++d) {
                const int i = r0 + d;
                const bool in = i >= sh.rz_lo && i <= sh.rz_hi;
                slot[d] = gbl(A.b.rrow)[sh.rrow_off + (in ? i : 0)];
                slot[d] = in ? slot[d] : -1;
            }
#pragma unroll
            for (int d = 0; d < FR; ++d) rix[d] = ld_sel(A.b.rmap, (long long)slot[d] * ld + x, slot[d] >= 0, -1);
        }
        if (!fastx) xuho = xh_o[(r0 + ud) * 2 * n1 + (cuh.col >= 0 ? cuh.col : 0)];
        sv = gbl(A.b.srcf)[sh.src_off + (has_src ? k : 0)];  // used only when has_src
    }

    // ---- stage the own rows (with halo) for the x chain
#pragma unroll
    for (int d = 0; d < FR; ++d) Ps[d][lane + 3] = p[d + 3];
#pragma unroll
    for (int q = 0; q < HP; ++q) Ps[xh.d[q]][xhalo_lds_col(xh, q, x0)] = hv[q];
    wave_sync();

    // ---- u nodes x0-1+c, c = 0..66 (LDS column c); c = lane, then c = 64..66 for all rows
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        if (EDGE && r0 + d >= pnz) break;
        const float *P = Ps[d];
        if (fastx) {
            Us[d][lane] = g.c0x * (P[lane + 2] - P[lane + 1]) + g.c1x * (P[lane + 3] - P[lane]);
        } else {
            float u = cu.A * (P[lane + 2] - P[lane + 1]) + cu.B * (P[lane + 3] - P[lane]);
            const float pn = cu.cb * xuo[d] + cu.ca * u;
            u = u + pn;
            Us[d][lane] = u;
            if (cu.col >= 0 && lane >= 1) xh_n[(r0 + d) * 2 * n1 + cu.col] = pn;  // owned: x0 <= j < x0+64
        }
    }
    {
        const float *P = Ps[ud];
        float u;
        if (fastx) {
            u = g.c0x * (P[uc + 2] - P[uc + 1]) + g.c1x * (P[uc + 3] - P[uc]);
        } else {
            u = cuh.A * (P[uc + 2] - P[uc + 1]) + cuh.B * (P[uc + 3] - P[uc]);
            const float pn = cuh.cb * xuho + cuh.ca * u;
            u = u + pn;
            if (cuh.col >= 0 && uc == 64) xh_n[(r0 + ud) * 2 * n1 + cuh.col] = pn;  // owned: node x0+63
        }
        Us[ud][uok ? uc : HJ] = u;
    }
    wave_sync();

    // ---- v nodes r0-1+n from the register rows (p[j] is p[n+2])
    float v[FVN];
#pragma unroll
    for (int n = 0; n < FVN; ++n) {
        const int j = r0 - 1 + n;
        float vv;
        if (zplain) {
            vv = g.c0z * (p[n + 2] - p[n + 1]) + g.c1z * (p[n + 3] - p[n]);
        } else {
            if (j < 0 || j >= pnz) vv = 0.f;
            else if (n == 1 && j == 0) vv = v_row0(p[n + 2], p[n + 3], p[n + 4], p[n + 5], g.iz840);
            else if (n == 2 && j == 1) vv = v_row1(p[n + 1], p[n + 2], p[n + 3], p[n + 4], g.iz840);
            else if (j == pnz - 1) vv = (p[n + 2] - p[n + 1]) * g.iz;
            else vv = g.c0z * (p[n + 2] - p[n + 1]) + g.c1z * (p[n + 3] - p[n]);
            if (j >= pnz - npml && j < pnz) {
                const int q = j - (pnz - npml);
                const float ps = TB(5, 2 * q) * zho[n] + TB(4, 2 * q) * vv;
                vv = vv + ps;
                if (n >= 1 && n <= FR) zh_n[q * ld + x] = ps;  // own nodes r0 .. r0+FR-1
            }
        }
        v[n] = vv;
    }

    // ---- output rows
    const int bx = x - npml, nx = g.nx, nz = g.nz;
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        const int i = r0 + d;
        if (EDGE && i >= pnz) break;
        const float *U = Us[d];
        float ux, vz;
        if (fastx) {
            ux = g.c0x * (U[lane + 2] - U[lane + 1]) + g.c1x * (U[lane + 3] - U[lane]);
        } else {
            ux = cx.A * (U[lane + 2] - U[lane + 1]) + cx.B * (U[lane + 3] - U[lane]);
            const float pn = cx.cb * xdo[d] + cx.ca * ux;
            ux = ux + pn;
            if (cx.col >= 0) xd_n[i * 2 * npml + cx.col] = pn;
        }
        if (zplain) {
            vz = g.c0z * (v[d + 2] - v[d + 1]) + g.c1z * (v[d + 3] - v[d]);
        } else {
            if (d == 0 && i == 0) vz = vz_row0(v[d + 1], v[d + 2], v[d + 3], v[d + 4], v[d + 5], g.iz24);  // nodes 0..4
            else if (i == pnz - 1) vz = (v[d + 2] - v[d + 1]) * g.iz;
            else vz = g.c0z * (v[d + 2] - v[d + 1]) + g.c1z * (v[d + 3] - v[d]);
            if (i >= pnz - npml) {
                const int q = i - (pnz - npml);
                const float ps = TB(5, 2 * q + 1) * zdo[d] + TB(4, 2 * q + 1) * vz;
                vz = vz + ps;
                zd_n[q * ld + x] = ps;
            }
        }
        if (has_src && i == sh.isz) ux = (x == sh.isx) ? ux + sv : ux;
        const float X = ux + vz;
        const float pnew = leapfrog(p[d + 3], po[d], pk[d] * X);
        if (MODE == FWI_FWD_XSEG) {
            // interior strips lie inside the physical grid (fwi_fwd_strip_interior)
            const int jx = x - npml;
            if (!EDGE || (jx >= 0 && jx < g.nx && i < g.nz))
                gbl(A.xseg)[(long long)(k - A.xseg_k0) * A.xseg_step + ((long long)s * g.nz + i) * g.nx + jx] = X;
        }
        if (!EDGE) {
            buf_st(pnew, rold, vo, i * rowb);
            continue;
        }
        if (x < pnx) buf_st(pnew, rold, vo, i * rowb);
        if (TRACES && has_recv && rix[d] >= 0)
            gbl(A.syn)[sh.trace_off + (long long)k * sh.nr + rix[d]] = pnew;
        if (MODE == FWI_FWD_FRAMES && (frame_rows || frame_cols) && bx >= 0 && bx < nx && i < nz &&
            (i < 3 || i >= nz - 3 || bx < 3 || bx >= nx - 3)) {
            const long long so2 = sh.strip_off + (long long)k * g.sl;
            if (i < 3 || i >= nz - 3) {
                const int pos = (i < 3) ? i * nx + bx : (3 + nz - 1 - i) * nx + bx;
                gbl(A.pstrip)[so2 + pos] = pnew;
                gbl(A.xstrip)[so2 + pos] = X;
            }
            if (bx < 3 || bx >= nx - 3) {
                const int pos = 6 * nx + i * 6 + ((bx < 3) ? bx : 3 + (nx - 1 - bx));
                gbl(A.pstrip)[so2 + pos] = pnew;
                gbl(A.xstrip)[so2 + pos] = X;
            }
        }
    }
}

// The parameter block lives in device memory (not in kernel arguments, which
// are all loaded at entry): each body then loads only the fields it uses, and
// the interior body keeps its scalar registers free of the general body's.
// The shot table and the tile list come as their own read-only kernel
// arguments so that they are read through the scalar cache (a pointer fetched
// from the parameter block would be read with vector loads, one memory
// latency per wave before its field loads can issue).  One instantiation
// runs the interior strips, the other the general strips: each list holds the
// logical strips (s nsz + sz) ntx + tx of its class (fwi_fwd_strip_interior),
// four consecutive entries per workgroup, one per wave.  A general strip whose
// special rules are all on the x side or all on the z side (85 % of them at
// 1024x320) runs the body compiled for just those rules: fewer instructions
// than the body with every rule (-4 % on the general launch).
template <bool EDGE, int MODE>
__global__ __launch_bounds__(256) void k_fwd_tile(const FwiFwdArgs *__restrict__ Ap,
                                                  const FwiShotDev *__restrict__ shots,
                                                  const int *__restrict__ strips, int nst, const int *kbase, int koff)
{
    static_assert(FR == FWI_STRIP, "one strip per wave");
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiFwdArgs &A = *Ap;
    const FwiGeomDev &g = A.g;
    // the wave index is uniform within a wave: keep it (and the rows) in SGPRs
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, (nst + 3) >> 2);
    if (lb < 0) return;
    __shared__ float Psh[4][FR][HC];
    __shared__ float Ush[4][FR][HC];
    __shared__ float stab[EDGE ? FWI_TABN : 1];
    if (EDGE) stage_tables_256(stab, A.b.tab, g.npml);  // the only workgroup barrier
    const int e = 4 * lb + w;
    if (e >= nst) return;
    const int nsz = (g.pnz + FR - 1) / FR, l = strips[e];
    const int tx = l % g.ntx, rest = l / g.ntx, sz = rest % nsz, s = rest / nsz;
    const FwiShotDev sh = shots[s];
    if (k >= sh.nt) return;
    if constexpr (EDGE) {
        const int c = fwi_fwd_strip_class(g, sh, tx * TX, sz * FR);
        if ((c & ~(FWI_CLS_X | FWI_CLS_FR)) == 0)
            fwd_tile_rows<EDGE, MODE, FWI_CLS_X | FWI_CLS_FR>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else if ((c & ~(FWI_CLS_Z | FWI_CLS_FR)) == 0)
            fwd_tile_rows<EDGE, MODE, FWI_CLS_Z | FWI_CLS_FR>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else
            fwd_tile_rows<EDGE, MODE>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    } else {
        fwd_tile_rows<EDGE, MODE>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    }
}

// The forward step as one launch: the interior and general strips' workgroups
// in one grid, each running its own body, so that no step waits on another
// stream.  work[b] = (role << 28) | workgroup index in the role's list (role
// 0: interior strips s_int, 1: general strips s_gen), -1: idle; the host
// orders the workgroups per XCD, the general ones spread through the
// interior ones (fwi_work_order).  The grid takes the general body's
// registers, so every strip runs at its occupancy.
template <int MODE>
__global__ __launch_bounds__(256) void k_fwd_mixed(const FwiFwdArgs *__restrict__ Ap,
                                                   const FwiShotDev *__restrict__ shots, const int *__restrict__ work,
                                                   const int *__restrict__ s_int, int n_int,
                                                   const int *__restrict__ s_gen, int n_gen, const int *kbase, int koff)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const int wv = work[blockIdx.x];
    if (wv < 0) return;
    const FwiFwdArgs &A = *Ap;
    const FwiGeomDev &g = A.g;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool gen = (wv >> 28) != 0;
    const int lb = wv & 0x0FFFFFFF;
    __shared__ float Psh[4][FR][HC];
    __shared__ float Ush[4][FR][HC];
    __shared__ float stab[FWI_TABN];
    if (gen) stage_tables_256(stab, A.b.tab, g.npml);  // uniform per workgroup: its only barrier
    const int e = 4 * lb + w;
    if (e >= (gen ? n_gen : n_int)) return;
    const int nsz = (g.pnz + FR - 1) / FR, l = (gen ? s_gen : s_int)[e];
    const int tx = l % g.ntx, rest = l / g.ntx, sz = rest % nsz, s = rest / nsz;
    const FwiShotDev sh = shots[s];
    if (k >= sh.nt) return;
    if (gen) {
        const int c = fwi_fwd_strip_class(g, sh, tx * TX, sz * FR);
        if ((c & ~(FWI_CLS_X | FWI_CLS_FR)) == 0)
            fwd_tile_rows<true, MODE, FWI_CLS_X | FWI_CLS_FR>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else if ((c & ~(FWI_CLS_Z | FWI_CLS_FR)) == 0)
            fwd_tile_rows<true, MODE, FWI_CLS_Z | FWI_CLS_FR>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else
            fwd_tile_rows<true, MODE>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    } else {
        fwd_tile_rows<false, MODE>(A, sh, stab, Psh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    }
}

// ============================================================================
// adjoint: exact transpose, x chain
// ============================================================================

// C^T of the ux-node CPML (backward in time) applied to w at p node q
__device__ __forceinline__ float adj_filt_d(const FwiGeomDev &g, const float *tab, int q, float w, int r,
                                            const float *xd_o, float *xd_n, bool own)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    if (q >= 0 && q < npml) {
        float c = xd_o[r * 2 * npml + q];
        c = TB(1, 2 * q + 1) * c + TB(0, 2 * q + 1) * w;
        w = w + c;
        if (own) xd_n[r * 2 * npml + q] = c;
    } else if (q >= pnx - npml && q < pnx) {
        const int m = q - (pnx - npml);
        float c = xd_o[r * 2 * npml + npml + m];
        c = TB(3, 2 * m + 1) * c + TB(2, 2 * m + 1) * w;
        w = w + c;
        if (own) xd_n[r * 2 * npml + npml + m] = c;
    }
    return w;
}

// rows of the second x derivative (p node i <- u nodes): taps u[i+1],u[i],u[i+2],u[i-1]
__device__ __forceinline__ float hxA(const XConst &g, int i)
{ return (i < 0 || i >= g.pnx) ? 0.f : ((i == 0 || i == g.pnx - 1) ? g.ix : g.c0x); }
__device__ __forceinline__ float hxB(const XConst &g, int i)
{ return (i < 0 || i >= g.pnx) ? 0.f : ((i == 0 || i == g.pnx - 1) ? -g.ix : -g.c0x); }
__device__ __forceinline__ float hxC(const XConst &g, int i)
{ return (i < 1 || i >= g.pnx - 1) ? 0.f : g.c1x; }
__device__ __forceinline__ float hxD(const XConst &g, int i)
{ return (i < 1 || i >= g.pnx - 1) ? 0.f : -g.c1x; }
// rows of the first x derivative (u node j <- p nodes): taps p[j],p[j-1],p[j+1],p[j-2]
__device__ __forceinline__ float gxE(const XConst &g, int j)
{ return (j <= 0 || j >= g.pnx) ? 0.f : ((j == 1 || j == g.pnx - 1) ? g.ix : g.c0x); }
__device__ __forceinline__ float gxF(const XConst &g, int j)
{ return (j <= 0 || j >= g.pnx) ? 0.f : ((j == 1 || j == g.pnx - 1) ? -g.ix : -g.c0x); }
__device__ __forceinline__ float gxG(const XConst &g, int j)
{ return (j <= 1 || j >= g.pnx - 1) ? 0.f : g.c1x; }
__device__ __forceinline__ float gxH(const XConst &g, int j)
{ return (j <= 1 || j >= g.pnx - 1) ? 0.f : -g.c1x; }

// H^T gather at u node j (wi = LDS index of w'[j]) followed by C^T of the u-node CPML
__device__ __forceinline__ float adj_u_node(const FwiGeomDev &g, const float *tab, const float *Ws, int j,
                                            int wi, int r, const float *xh_o, float *xh_n, int x0)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1, n1 = npml + 1;
    if (j < 0 || j > pnx) return 0.f;
    const XConst xk = xconst(g);
    float u = hxA(xk, j - 1) * Ws[wi - 1] + hxB(xk, j) * Ws[wi] + hxC(xk, j - 2) * Ws[wi - 2] + hxD(xk, j + 1) * Ws[wi + 1];
    const bool own = (j >= x0) && (j < x0 + TX);
    if (j <= npml) {
        float c = xh_o[r * 2 * n1 + j];
        c = TB(1, 2 * j) * c + TB(0, 2 * j) * u;
        u = u + c;
        if (own) xh_n[r * 2 * n1 + j] = c;
    }
    if (j >= pnx - npml) {
        const int q = j - (pnx - npml);
        float c = xh_o[r * 2 * n1 + n1 + q];
        c = TB(3, 2 * q) * c + TB(2, 2 * q) * u;
        u = u + c;
        if (own) xh_n[r * 2 * n1 + n1 + q] = c;
    }
    return u;
}

// C^T of the ux-node CPML at p node q from the old memory co (new value in
// cn): the arithmetic of adj_filt_d, memory traffic left to the caller
__device__ __forceinline__ float adj_filt_edge(const FwiGeomDev &g, const float *tab, int q, float w, float co,
                                               float &cn)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    cn = 0.f;
    if (q >= 0 && q < npml) {
        cn = TB(1, 2 * q + 1) * co + TB(0, 2 * q + 1) * w;
        w = w + cn;
    } else if (q >= pnx - npml && q < pnx) {
        const int m = q - (pnx - npml);
        cn = TB(3, 2 * m + 1) * co + TB(2, 2 * m + 1) * w;
        w = w + cn;
    }
    return w;
}
// H^T gather at u node j then C^T of the u-node CPML: the arithmetic of adj_u_node
__device__ __forceinline__ float adj_u_edge(const FwiGeomDev &g, const float *tab, const float *Ws, int j, int wi,
                                            float co, float &cn)
{
    const int pnx = g.pnx, npml = g.npml, T = 2 * npml + 1;
    cn = 0.f;
    if (j < 0 || j > pnx) return 0.f;
    const XConst xk = xconst(g);
    float u = hxA(xk, j - 1) * Ws[wi - 1] + hxB(xk, j) * Ws[wi] + hxC(xk, j - 2) * Ws[wi - 2] + hxD(xk, j + 1) * Ws[wi + 1];
    if (j <= npml) {
        cn = TB(1, 2 * j) * co + TB(0, 2 * j) * u;
        u = u + cn;
    } else if (j >= pnx - npml) {
        const int q = j - (pnx - npml);
        cn = TB(3, 2 * q) * co + TB(2, 2 * q) * u;
        u = u + cn;
    }
    return u;
}

template <bool FAST, int N>
__device__ __forceinline__ void adj_xblock(const FwiGeomDev &g, const float *tab, const float *bs, const float *bh,
                                           float (*Ws)[72], float (*Us)[72], int x0, int lane, int rb, int z0,
                                           int z1, const float *xh_o, float *xh_n, const float *xd_o, float *xd_n,
                                           float *box)
{
    const int hq = halo_col(x0, lane);
    const int hidx = (lane < 3) ? lane : lane + 64;
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        if (r >= z0 && r < z1) {
            float wm = bs[d];
            if (!FAST) wm = adj_filt_d(g, tab, x0 + lane, wm, r, xd_o, xd_n, x0 + lane < g.pnx);
            Ws[d][lane + 3] = wm;
            if (lane < 6) {
                float wh = bh[d];
                if (!FAST) wh = adj_filt_d(g, tab, hq, wh, r, xd_o, xd_n, false);
                Ws[d][hidx] = wh;
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        if (r >= z0 && r < z1) {
            const float *W = Ws[d];
            float *U = Us[d];
            if (FAST) {
                U[lane] = g.c0x * (W[lane + 1] - W[lane + 2]) + g.c1x * (W[lane] - W[lane + 3]);
                if (lane < 3)
                    U[64 + lane] = g.c0x * (W[lane + 65] - W[lane + 66]) + g.c1x * (W[lane + 64] - W[lane + 67]);
            } else {
                U[lane] = adj_u_node(g, tab, W, x0 - 1 + lane, lane + 2, r, xh_o, xh_n, x0);
                if (lane < 3) U[64 + lane] = adj_u_node(g, tab, W, x0 + 63 + lane, lane + 66, r, xh_o, xh_n, x0);
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int d = 0; d < N; ++d) {
        const int r = rb + d;
        box[d] = 0.f;
        if (r >= z0 && r < z1) {
            const float *U = Us[d];
            if (FAST) {
                box[d] = g.c0x * (U[lane + 1] - U[lane + 2]) + g.c1x * (U[lane] - U[lane + 3]);
            } else {
                const int i = x0 + lane, ui = lane + 1;
                box[d] = (i >= g.pnx) ? 0.f
                                      : gxE(xconst(g), i) * U[ui] + gxF(xconst(g), i + 1) * U[ui + 1] +
                                            gxG(xconst(g), i - 1) * U[ui - 1] + gxH(xconst(g), i + 2) * U[ui + 2];
            }
        }
    }
}

__global__ __launch_bounds__(FWI_TX * FWI_MAXSG) void k_adj_step(FwiAdjArgs A, const int *kbase, int koff)
{
    // step index: koff alone, or a device counter plus koff inside a captured graph
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiGeomDev &g = A.g;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ntiles = g.ntx * g.ntz;
    const int lb = xcd_remap(blockIdx.x, ntiles * A.b.ngroups);
    if (lb < 0) return;
    const int grp = lb / ntiles, t = lb - grp * ntiles;
    const int tzi = t / g.ntx, txi = t - tzi * g.ntx;
    const int s = grp * A.b.sg + wv;
    __shared__ float stab[FWI_TABN];
    stage_tables(stab, A.b.tab, g.npml);
    if (s >= A.b.nshots) return;
    const FwiShotDev sh = A.b.shots[s];
    if (k >= sh.nt) return;

    __shared__ float lds[FWI_MAXSG][2][RB][72];
    float(*Ws)[72] = lds[wv][0];
    float(*Us)[72] = lds[wv][1];

    const int npml = g.npml, pnx = g.pnx, pnz = g.pnz, ld = g.ld, T = 2 * npml + 1;
    const float *tab = stab;
    const int x0 = txi * TX, x = x0 + lane;
    const int z0 = tzi * g.tz, z1 = min(z0 + g.tz, pnz);
    const long long so = (long long)s * A.b.fs;
    // mu^{k+2} in M[k&1] (read), mu^{k+3} in M[(k+1)&1] overwritten by mu^{k+1}
    const int cur = k & 1, prv = cur ^ 1;
    const float *mnow = (cur ? A.M1 : A.M0) + so;
    float *mold = (cur ? A.M0 : A.M1) + so;
    const float *ekp = A.b.ekp;
    // chi after step k goes to buffer (k & 1); it was read from buffer ((k+1) & 1)
    const float *xh_o = A.xh + s * A.b.xh_s + prv * A.b.xh_b;
    float *xh_n = A.xh + s * A.b.xh_s + cur * A.b.xh_b;
    const float *xd_o = A.xd + s * A.b.xd_s + prv * A.b.xd_b;
    float *xd_n = A.xd + s * A.b.xd_s + cur * A.b.xd_b;
    const float *zh_o = A.zh + s * A.b.zh_s + prv * A.b.zh_b;
    float *zh_n = A.zh + s * A.b.zh_s + cur * A.b.zh_b;
    const float *zd_o = A.zd + s * A.b.zd_s + prv * A.b.zd_b;
    float *zd_n = A.zd + s * A.b.zd_s + cur * A.b.zd_b;
    const bool fastx = (x0 >= npml + 3) && (x0 + 66 <= pnx - npml - 1);
    const float c0z = g.c0z, c1z = g.c1z, izf = g.iz;
    // transposed free-surface rows
    const double i24 = (double)g.iz24, i840 = (double)g.iz840;
    const float hz0[5] = {(float)(-22. * i24), (float)(17. * i24), (float)(9. * i24), (float)(-5. * i24),
                          (float)(1. * i24)};
    const float gz0[4] = {(float)(3675.0 * i840), (float)(-1225.0 * i840), (float)(441.0 * i840),
                          (float)(-75.0 * i840)};
    const float gz1[4] = {(float)(-1085.0 * i840), (float)(1015.0 * i840), (float)(-63.0 * i840),
                          (float)(5.0 * i840)};

    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;  // w'_z[r-3..r] (z-filtered mu^{k+2})
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;  // raw mu^{k+2}[r-3..r]
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;  // v~'[r-4..r-1]
    float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;  // out_x[r-3..r]
    float wz0 = 0.f, tv0 = 0.f, tv1 = 0.f;          // top rows held by the first tile

    const int rend = z1 + 2;
    const int hq = halo_col(x0, lane);
    auto fetch = [&](int r, float &fs, float &fh, float &fo, float &fk) {
        fs = fh = fo = fk = 0.f;
        if (r > rend) return;
        if (r >= 0 && r < pnz) fs = mnow[(long long)r * ld + x];
        if (r >= z0 && r < z1 && hq >= 0 && hq < pnx) fh = mnow[(long long)r * ld + hq];
        const int io = r - 3;
        if (io >= z0 && io < z1) {
            fo = mold[(long long)io * ld + x];
            fk = ekp[(long long)io * ld + x];
        }
    };
    auto body = [&](int r, float wr, float oxr, float mold_i, float ekp_i) {
        // z chain: C^T of the vz-node CPML
        float wz = wr;
        if (r >= pnz - npml && r < pnz) {
            const int q = r - (pnz - npml);
            float c = zd_o[q * ld + x];
            c = TB(5, 2 * q + 1) * c + TB(4, 2 * q + 1) * wz;
            wz = wz + c;
            if (r >= z0 && r < z1) zd_n[q * ld + x] = c;
        }
        if (r == 0) wz0 = wz;
        a0 = a1; a1 = a2; a2 = a3; a3 = wz;
        m0 = m1; m1 = m2; m2 = m3; m3 = wr;
        q0 = q1; q1 = q2; q2 = q3; q3 = oxr;

        // v~ at half node j = r - 1 from w'[j-2..j+1] = a0..a3  (H_z^T)
        const int j = r - 1;
        float tv = 0.f;
        if (j >= 0 && j <= pnz) {
            if (j > 4 && j + 1 <= pnz - 2) {
                tv = c0z * (a1 - a2) + c1z * (a0 - a3);
            } else {
                // interior rows i in [1, pnz-1): taps v[i+1]=c0, v[i]=-c0, v[i+2]=c1, v[i-1]=-c1
                if (j - 1 >= 1 && j - 1 < pnz - 1) tv += c0z * a1;
                if (j >= 1 && j < pnz - 1) tv -= c0z * a2;
                if (j - 2 >= 1 && j - 2 < pnz - 1) tv += c1z * a0;
                if (j + 1 >= 1 && j + 1 < pnz - 1) tv -= c1z * a3;
                if (j <= 4) tv += hz0[j] * wz0;
                if (j == pnz) tv += izf * a1;
                if (j == pnz - 1) tv -= izf * a2;
            }
            if (j >= pnz - npml && j < pnz) {
                const int q = j - (pnz - npml);
                float c = zh_o[q * ld + x];
                c = TB(5, 2 * q) * c + TB(4, 2 * q) * tv;
                tv = tv + c;
                if (j >= z0 && j < z1) zh_n[q * ld + x] = c;
            }
        }
        if (j == 0) tv0 = tv;
        if (j == 1) tv1 = tv;
        b0 = b1; b1 = b2; b2 = b3; b3 = tv;

        const int i = r - 3;  // output row: v~'[i-1..i+2] = b0..b3
        if (i >= z0 && i < z1) {
            float oz = 0.f;
            if (i > 3 && i - 1 >= 2 && i + 2 < pnz - 1) {
                oz = c0z * (b1 - b2) + c1z * (b0 - b3);
            } else {
                // rows j in [2, pnz-1): taps p[j]=c0, p[j-1]=-c0, p[j+1]=c1, p[j-2]=-c1
                if (i >= 2 && i < pnz - 1) oz += c0z * b1;
                if (i + 1 >= 2 && i + 1 < pnz - 1) oz -= c0z * b2;
                if (i - 1 >= 2 && i - 1 < pnz - 1) oz += c1z * b0;
                if (i + 2 >= 2 && i + 2 < pnz - 1) oz -= c1z * b3;
                if (i <= 3) oz += gz0[i] * tv0 + gz1[i] * tv1;
                if (i == pnz - 1) oz += izf * b1;
                if (i == pnz - 2) oz -= izf * b2;
            }
            float out = q0 + oz;
            if (i >= sh.rz_lo && i <= sh.rz_hi) {
                const int slot = gbl(A.b.rrow)[sh.rrow_off + i];
                if (slot >= 0) {
                    const int ri = gbl(A.b.rmap)[slot * ld + x];
                    if (ri >= 0) out = out + gbl(A.res)[sh.trace_off + (long long)k * sh.nr + ri];
                }
            }
            const long long o = (long long)i * ld + x;
            const float mnew = leapfrog(m0, mold_i, ekp_i * out);
            if (x < pnx) mold[o] = mnew;
        }
    };
    // interior rows: fast transposed stencils, no strip and no receiver
    auto plain_body = [&](int r, float wr, float oxr, float mold_i, float ekp_i) {
        a0 = a1; a1 = a2; a2 = a3; a3 = wr;
        m0 = m1; m1 = m2; m2 = m3; m3 = wr;
        q0 = q1; q1 = q2; q2 = q3; q3 = oxr;
        const float tv = c0z * (a1 - a2) + c1z * (a0 - a3);
        b0 = b1; b1 = b2; b2 = b3; b3 = tv;
        const float oz = c0z * (b1 - b2) + c1z * (b0 - b3);
        const float out = q0 + oz;
        mold[(r - 3) * ld + x] = leapfrog(m0, mold_i, ekp_i * out);
    };
    auto gen_block = [&](auto Nc, int rb) {
        constexpr int N = decltype(Nc)::value;
        float bs[N], bh[N], bo[N], bk[N], bu[N];
#pragma unroll
        for (int d = 0; d < N; ++d) fetch(rb + d, bs[d], bh[d], bo[d], bk[d]);
        if (rb + N - 1 >= z0 && rb < z1) {
            if (fastx) adj_xblock<true, N>(g, tab, bs, bh, Ws, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
            else adj_xblock<false, N>(g, tab, bs, bh, Ws, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
        } else {
#pragma unroll
            for (int d = 0; d < N; ++d) bu[d] = 0.f;
        }
#pragma unroll
        for (int d = 0; d < N; ++d)
            if (rb + d <= rend) body(rb + d, bs[d], bu[d], bo[d], bk[d]);
    };
    auto run_block = [&](int rb, const float *bs, const float *bh, const float *bo, const float *bk) {
        float bu[RB];
        const int ilo = rb - 3, ihi = rb + RB - 4;
        const bool plain = fastx && ilo >= 5 && ihi < g.nz - 3 && (sh.rz_hi < ilo || sh.rz_lo > ihi);
        if (plain) {
            adj_xblock<true, RB>(g, tab, bs, bh, Ws, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
#pragma unroll
            for (int d = 0; d < RB; ++d) plain_body(rb + d, bs[d], bu[d], bo[d], bk[d]);
        } else {
            if (fastx) adj_xblock<true, RB>(g, tab, bs, bh, Ws, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
            else adj_xblock<false, RB>(g, tab, bs, bh, Ws, Us, x0, lane, rb, z0, z1, xh_o, xh_n, xd_o, xd_n, bu);
#pragma unroll
            for (int d = 0; d < RB; ++d) body(rb + d, bs[d], bu[d], bo[d], bk[d]);
        }
    };
    using I3 = std::integral_constant<int, 3>;
    gen_block(I3{}, z0 - 3);
    gen_block(I3{}, z0);
    int rb = z0 + 3;
    const int rmain = z1 - 1;
    if (rb + RB - 1 <= rmain) {
        float As[RB], Ah[RB], Ao[RB], Ak[RB], Bs[RB], Bh[RB], Bo[RB], Bk[RB];
#pragma unroll
        for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ao[d], Ak[d]);
        for (;;) {
            const bool nb = rb + 2 * RB - 1 <= rmain;
            if (nb && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, Bs[d], Bh[d], Bo[d], Bk[d]);
            }
            run_block(rb, As, Ah, Ao, Ak);
            rb += RB;
            if (!nb) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, Bs[d], Bh[d], Bo[d], Bk[d]);
            }
            const bool na = rb + 2 * RB - 1 <= rmain;
            if (na && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, As[d], Ah[d], Ao[d], Ak[d]);
            }
            run_block(rb, Bs, Bh, Bo, Bk);
            rb += RB;
            if (!na) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ao[d], Ak[d]);
            }
        }
    }
    for (; rb <= rend; rb += 3) gen_block(I3{}, rb);
}

// ============================================================================
// adjoint, 2-D tile form: the exact transpose of k_fwd_tile's step with the
// same tiling (4 waves x FR contiguous rows of one shot per workgroup).
// mu^{k+2} rows r0-3 .. r0+FR+2 sit in registers; the z chain (C_d^T, H_z^T,
// C_h^T, G_z^T with the transposed free-surface rows) runs in registers, the
// x chain (C_d^T on the staged row, H_x^T + C_h^T at u nodes, G_x^T) through
// wave-private LDS rows.  Same arithmetic and CPML ownership as k_adj_step.
// ============================================================================
// FUSED: ekp of the rows comes from the caller (ek) and mu^{k+1} of the
// rows is also left in keep (rows past the grid untouched), for the fused
// backward step.
template <bool EDGE, bool FUSED = false, int CLS = FWI_CLS_ALL>
__device__ __forceinline__ void adj_tile_rows(const FwiAdjArgs &A, const FwiShotDev &sh, const float *tab,
                                              float (*Ws)[HC], float (*Us)[HC], int s, int k, int x0, int r0,
                                              int lane, const float *ek = nullptr, float *keep = nullptr)
{
    const FwiGeomDev &g = A.g;
    const int npml = g.npml, pnx = g.pnx, pnz = g.pnz, ld = g.ld, T = 2 * npml + 1;
    const int x = x0 + lane;
    const long long so = (long long)s * A.b.fs;
    // mu^{k+2} in M[k&1] (read), mu^{k+3} in M[(k+1)&1] overwritten by mu^{k+1}
    const int cur = k & 1, prv = cur ^ 1;
    const float *mnow = (cur ? A.M1 : A.M0) + so;
    float *mold = (cur ? A.M0 : A.M1) + so;
    const unsigned fbytes = (unsigned)(A.b.fs * sizeof(float));
    const brsrc rnow = buf_rsrc(mnow, fbytes), rold = buf_rsrc(mold, fbytes), rk = buf_rsrc(A.b.ekp, fbytes);
    const int vo = x * 4, rowb = ld * 4;

    // ---- every load of the wave
    float m[FPR], mo[FR], mk[FR];
#pragma unroll
    for (int j = 0; j < FPR; ++j) {
        const int zr = r0 - 3 + j;
        m[j] = EDGE ? buf_ld_row(rnow, vo, zr, pnz, rowb) : buf_ld(rnow, vo, zr * rowb);
    }
    XHalo<HP> xh;
    xhalo_plan(xh, lane, x0, FR);
    float hv[HP];
    bool hok[HP];
#pragma unroll
    for (int q = 0; q < HP; ++q) {
        hok[q] = xh.in[q];
        if (EDGE) hok[q] = hok[q] && r0 + xh.d[q] < pnz && xh.col[q] >= 0 && xh.col[q] < pnx;
        hv[q] = buf_ld(rnow, hok[q] ? (xh.d[q] * ld + xh.col[q]) * 4 : vo, r0 * rowb);
        hv[q] = hok[q] ? hv[q] : 0.f;
    }
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        const int i = r0 + d;
        mo[d] = EDGE ? buf_ld_row(rold, vo, i, pnz, rowb) : buf_ld(rold, vo, i * rowb);
        if (FUSED) mk[d] = ek[d];
        else mk[d] = EDGE ? buf_ld_row(rk, vo, i, pnz, rowb) : buf_ld(rk, vo, i * rowb);
    }
    float m0 = 0.f;  // row 0 (strips shorter than 6 rows: the free-surface rule below)
    if constexpr (FR <= 5)
        if (EDGE) m0 = buf_ld(rnow, vo, 0);

    // chi after step k goes to buffer (k & 1); it was read from buffer ((k+1) & 1).
    // One buffer resource per array over both buffers of the shot: the lane's
    // column in the byte offset (fixed for the wave), the row and the buffer
    // in the uniform offset, so no 64-bit address arithmetic per row.
    const brsrc rxh = buf_rsrc(A.xh + s * A.b.xh_s, (unsigned)(8 * A.b.xh_b));
    const brsrc rxd = buf_rsrc(A.xd + s * A.b.xd_s, (unsigned)(8 * A.b.xd_b));
    const brsrc rzh = buf_rsrc(A.zh + s * A.b.zh_s, (unsigned)(8 * A.b.zh_b));
    const brsrc rzd = buf_rsrc(A.zd + s * A.b.zd_s, (unsigned)(8 * A.b.zd_b));
    const int xh_o = (int)(prv * A.b.xh_b) * 4, xh_n = (int)(cur * A.b.xh_b) * 4;  // byte offsets of the buffers
    const int xd_o = (int)(prv * A.b.xd_b) * 4, xd_n = (int)(cur * A.b.xd_b) * 4;
    const int zh_o = (int)(prv * A.b.zh_b) * 4, zh_n = (int)(cur * A.b.zh_b) * 4;
    const int zd_o = (int)(prv * A.b.zd_b) * 4, zd_n = (int)(cur * A.b.zd_b) * 4;

    // columns x0-3 .. x0+66 clear of the x CPML and of the grid edges: compact
    // x stencils (the form k_adj_step uses on such tiles)
    // (each special rule is compiled in only when the body's mask CLS allows it)
    const bool fastx = !EDGE || !(CLS & FWI_CLS_X) || ((x0 >= npml + 3) && (x0 + 66 <= pnx - npml - 1));
    const int ud0 = lane / 3;  // the extra-u-node pass: lane = 3 ud + (uc - 64)
    bool uok = lane < 3 * FR;
    if (EDGE) uok = uok && r0 + ud0 < pnz;
    const int ud = uok ? ud0 : 0, uc = uok ? 64 + lane - 3 * ud0 : 65;

    // ---- general tiles: every other load before the first store (vector memory
    // operations retire in order: a load issued after a store waits for it)
    float xdo[FR], xdho[HP], xuo[FR], xuho = 0.f, zdo[FPR], zho[FVN], rsv[FR];
    bool rhas[FR];
    // x strips, branch free: per-node coefficients computed once per wave (the
    // arithmetic of adj_filt_edge, adj_u_edge and the G_x^T gather; the CPML
    // coefficients are those of the forward's nodes, fwd_u_coef / fwd_ux_coef)
    XCoef cf, cu, cuh, cfh[HP];
    float hA = 0.f, hB = 0.f, hC = 0.f, hD = 0.f, hhA = 0.f, hhB = 0.f, hhC = 0.f, hhD = 0.f;
    float gE = 0.f, gF = 0.f, gG = 0.f, gH = 0.f;
    const int n1 = npml + 1;
    // which special rules can occur in this strip at all (uniform)
    const bool zplain = !EDGE || !(CLS & FWI_CLS_Z) || (r0 >= 6 && r0 + FR + 2 < pnz - npml);  // plain z chain
    const bool has_recv = EDGE && (CLS & FWI_CLS_RS) && sh.rz_lo <= sh.rz_hi && sh.rz_lo < r0 + FR && sh.rz_hi >= r0;
    if (EDGE && !fastx) {
        const XConst xk = xconst(g);
        cf = fwd_ux_coef(xk, tab, x);
#pragma unroll
        for (int q = 0; q < HP; ++q) cfh[q] = fwd_ux_coef(xk, tab, xh.col[q]);
        const int j = x0 - 1 + lane, jh = x0 - 1 + uc;
        cu = fwd_u_coef(xk, tab, j);
        cuh = fwd_u_coef(xk, tab, jh);
        hA = hxA(xk, j - 1); hB = hxB(xk, j); hC = hxC(xk, j - 2); hD = hxD(xk, j + 1);
        hhA = hxA(xk, jh - 1); hhB = hxB(xk, jh); hhC = hxC(xk, jh - 2); hhD = hxD(xk, jh + 1);
        gE = gxE(xk, x); gF = gxF(xk, x + 1); gG = gxG(xk, x - 1); gH = gxH(xk, x + 2);
    }
    if (EDGE) {
#pragma unroll
        for (int d = 0; d < FR; ++d) {
            // x strips: per lane, clamped index, no divergence; rows past the
            // grid read the last row (their nodes are not evaluated)
            const int i = min(r0 + d, pnz - 1);
            xdo[d] = xuo[d] = 0.f;
            if (!fastx) {
                xdo[d] = buf_ld(rxd, (cf.col >= 0 ? cf.col : 0) * 4, xd_o + i * 2 * npml * 4);
                xuo[d] = buf_ld(rxh, (cu.col >= 0 ? cu.col : 0) * 4, xh_o + i * 2 * n1 * 4);
            }
        }
#pragma unroll
        for (int d = 0; d < FR; ++d) {
            rhas[d] = false;
            rsv[d] = 0.f;
        }
        if (has_recv) {  // every row slot first, then the receiver indices, then the samples
            int slot[FR], rix[FR];
#pragma unroll
            for (int d = 0; d < FR; ++d) {
                const int i = r0 + d;
                const bool in = i >= sh.rz_lo && i <= sh.rz_hi;
                slot[d] = gbl(A.b.rrow)[sh.rrow_off + (in ? i : 0)];
                slot[d] = in ? slot[d] : -1;
            }
#pragma unroll
            for (int d = 0; d < FR; ++d) rix[d] = ld_sel(A.b.rmap, (long long)slot[d] * ld + x, slot[d] >= 0, -1);
#pragma unroll
            for (int d = 0; d < FR; ++d) {
                rhas[d] = rix[d] >= 0;
                rsv[d] = ld_sel(A.res, sh.trace_off + (long long)k * sh.nr + (rix[d] >= 0 ? rix[d] : 0),
                                slot[d] >= 0, 0.f);
            }
        }
#pragma unroll
        for (int q = 0; q < HP; ++q) {
            xdho[q] = 0.f;
            if (!fastx)
                xdho[q] = buf_ld(rxd, (xh.d[q] * 2 * npml + ((hok[q] && cfh[q].col >= 0) ? cfh[q].col : 0)) * 4,
                                 xd_o + r0 * 2 * npml * 4);
        }
        if (!fastx) xuho = buf_ld(rxh, (ud * 2 * n1 + (cuh.col >= 0 ? cuh.col : 0)) * 4, xh_o + r0 * 2 * n1 * 4);
        if (!zplain) {
#pragma unroll
            for (int j = 0; j < FPR; ++j) {
                const int q = r0 - 3 + j - (pnz - npml);
                zdo[j] = buf_ld_row(rzd, x * 4, q, npml, ld * 4, zd_o);
            }
#pragma unroll
            for (int n = 0; n < FVN; ++n) {
                const int q = r0 - 1 + n - (pnz - npml);
                zho[n] = buf_ld_row(rzh, x * 4, q, npml, ld * 4, zh_o);
            }
        }
    }

    // ---- x chain, step 1: w' = C_d^T mu on the own rows and the halo nodes
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        float wm = m[d + 3];
        if (!fastx && r0 + d < pnz) {
            const float cn = cf.cb * xdo[d] + cf.ca * wm;
            wm = wm + cn;
            if (cf.col >= 0) buf_st(cn, rxd, cf.col * 4, xd_n + (r0 + d) * 2 * npml * 4);  // x < pnx whenever col >= 0
        }
        Ws[d][lane + 3] = wm;
    }
#pragma unroll
    for (int q = 0; q < HP; ++q) {
        float w = hv[q];
        if (!fastx && hok[q]) w = w + (cfh[q].cb * xdho[q] + cfh[q].ca * w);  // halo node: owned by the neighbour
        Ws[xh.d[q]][xhalo_lds_col(xh, q, x0)] = w;
    }
    wave_sync();
    // ---- step 2: H_x^T then C_h^T at u nodes x0-1+c (LDS column c)
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        if (EDGE && r0 + d >= pnz) break;
        const float *W = Ws[d];
        if (fastx) {
            Us[d][lane] = g.c0x * (W[lane + 1] - W[lane + 2]) + g.c1x * (W[lane] - W[lane + 3]);
        } else {
            float u = hA * W[lane + 1] + hB * W[lane + 2] + hC * W[lane] + hD * W[lane + 3];
            const float cn = cu.cb * xuo[d] + cu.ca * u;
            u = u + cn;
            Us[d][lane] = u;
            if (cu.col >= 0 && lane >= 1) buf_st(cn, rxh, cu.col * 4, xh_n + (r0 + d) * 2 * n1 * 4);  // owned: x0 <= j < x0+64
        }
    }
    {
        const float *W = Ws[ud];
        float u;
        if (fastx) {
            u = g.c0x * (W[uc + 1] - W[uc + 2]) + g.c1x * (W[uc] - W[uc + 3]);
        } else {
            u = hhA * W[uc + 1] + hhB * W[uc + 2] + hhC * W[uc] + hhD * W[uc + 3];
            const float cn = cuh.cb * xuho + cuh.ca * u;
            u = u + cn;
            if (cuh.col >= 0 && uc == 64) buf_st(cn, rxh, (ud * 2 * n1 + cuh.col) * 4, xh_n + r0 * 2 * n1 * 4);  // owned: node x0+63
        }
        Us[ud][uok ? uc : HJ] = u;
    }
    wave_sync();

    // ---- z chain in registers: a = C_d^T mu (rows r0-3+j), tv = C_h^T H_z^T a (nodes r0-1+n)
    const float c0z = g.c0z, c1z = g.c1z, izf = g.iz;
    float a[FPR];
#pragma unroll
    for (int j = 0; j < FPR; ++j) {
        const int r = r0 - 3 + j;
        float wz = m[j];
        if (!zplain && r >= pnz - npml && r < pnz) {
            const int q = r - (pnz - npml);
            const float c = TB(5, 2 * q + 1) * zdo[j] + TB(4, 2 * q + 1) * wz;
            wz = wz + c;
            if (j >= 3 && j < FR + 3) buf_st(c, rzd, x * 4, zd_n + q * ld * 4);  // own rows
        }
        a[j] = wz;
    }
    float tv[FVN];
#pragma unroll
    for (int n = 0; n < FVN; ++n) {
        const int jn = r0 - 1 + n;
        float t = 0.f;
        if (zplain) {
            t = c0z * (a[n + 1] - a[n + 2]) + c1z * (a[n] - a[n + 3]);
        } else if (jn >= 0 && jn <= pnz) {
            if (jn > 4 && jn + 1 <= pnz - 2) {
                t = c0z * (a[n + 1] - a[n + 2]) + c1z * (a[n] - a[n + 3]);
            } else {
                // rows i in [1, pnz-1) of vz: taps v[i+1]=c0, v[i]=-c0, v[i+2]=c1, v[i-1]=-c1
                if (jn - 1 >= 1 && jn - 1 < pnz - 1) t += c0z * a[n + 1];
                if (jn >= 1 && jn < pnz - 1) t -= c0z * a[n + 2];
                if (jn - 2 >= 1 && jn - 2 < pnz - 1) t += c1z * a[n];
                if (jn + 1 >= 1 && jn + 1 < pnz - 1) t -= c1z * a[n + 3];
                // vz row 0 (free surface) reads v nodes 0..4, so those nodes
                // take a term of row 0: on the strip at r0 = 0, and with
                // strips of fewer than 6 rows on the next one too (row 0 is
                // then outside a[], read as m0: no z CPML at row 0)
                if (r0 <= 5 && jn >= 0 && jn <= 4) {
                    const double i24 = (double)g.iz24;
                    const double w0 = (jn == 0) ? -22. : (jn == 1) ? 17. : (jn == 2) ? 9. : (jn == 3) ? -5. : 1.;
                    t += (float)(w0 * i24) * (r0 == 0 ? a[3] : m0);
                }
                if (jn == pnz) t += izf * a[n + 1];
                if (jn == pnz - 1) t -= izf * a[n + 2];
            }
            if (jn >= pnz - npml && jn < pnz) {
                const int q = jn - (pnz - npml);
                const float c = TB(5, 2 * q) * zho[n] + TB(4, 2 * q) * t;
                t = t + c;
                if (n >= 1 && n <= FR) buf_st(c, rzh, x * 4, zh_n + q * ld * 4);  // own nodes r0 .. r0+FR-1
            }
        }
        tv[n] = t;
    }

    // ---- output rows: out = G_x^T u~ + G_z^T tv (+ residual), then the leapfrog
#pragma unroll
    for (int d = 0; d < FR; ++d) {
        const int i = r0 + d;
        if (EDGE && i >= pnz) break;
        const float *U = Us[d];
        float ox, oz;
        if (fastx) {
            ox = g.c0x * (U[lane + 1] - U[lane + 2]) + g.c1x * (U[lane] - U[lane + 3]);
        } else {
            ox = gE * U[lane + 1] + gF * U[lane + 2] + gG * U[lane] + gH * U[lane + 3];
        }
        if (zplain) {
            oz = c0z * (tv[d + 1] - tv[d + 2]) + c1z * (tv[d] - tv[d + 3]);
        } else {
            if (i > 3 && i - 1 >= 2 && i + 2 < pnz - 1) {
                oz = c0z * (tv[d + 1] - tv[d + 2]) + c1z * (tv[d] - tv[d + 3]);
            } else {
                // rows j in [2, pnz-1) of v: taps p[j]=c0, p[j-1]=-c0, p[j+1]=c1, p[j-2]=-c1
                oz = 0.f;
                if (i >= 2 && i < pnz - 1) oz += c0z * tv[d + 1];
                if (i + 1 >= 2 && i + 1 < pnz - 1) oz -= c0z * tv[d + 2];
                if (i - 1 >= 2 && i - 1 < pnz - 1) oz += c1z * tv[d];
                if (i + 2 >= 2 && i + 2 < pnz - 1) oz -= c1z * tv[d + 3];
                if (r0 == 0 && d <= 3) {  // v rows 0 and 1 (free surface) read p rows 0..3
                    const double i840 = (double)g.iz840;
                    const double w0 = (d == 0) ? 3675.0 : (d == 1) ? -1225.0 : (d == 2) ? 441.0 : -75.0;
                    const double w1 = (d == 0) ? -1085.0 : (d == 1) ? 1015.0 : (d == 2) ? -63.0 : 5.0;
                    oz += (float)(w0 * i840) * tv[1] + (float)(w1 * i840) * tv[2];
                }
                if (i == pnz - 1) oz += izf * tv[d + 1];
                if (i == pnz - 2) oz -= izf * tv[d + 2];
            }
        }
        float out = ox + oz;
        if (has_recv && rhas[d]) out = out + rsv[d];
        const float mnew = leapfrog(m[d + 3], mo[d], mk[d] * out);
        if (!EDGE || x < pnx) buf_st(mnew, rold, vo, i * rowb);
        if (FUSED) keep[d] = mnew;
    }
}


// interior / general strips of the lists built with fwi_adj_strip_interior
template <bool EDGE>
__global__ __launch_bounds__(256) void k_adj_tile(const FwiAdjArgs *__restrict__ Ap,
                                                  const FwiShotDev *__restrict__ shots,
                                                  const int *__restrict__ strips, int nst, const int *kbase, int koff)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiAdjArgs &A = *Ap;
    const FwiGeomDev &g = A.g;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, (nst + 3) >> 2);
    if (lb < 0) return;
    __shared__ float Wsh[4][FR][HC];
    __shared__ float Ush[4][FR][HC];
    __shared__ float stab[EDGE ? FWI_TABN : 1];
    if (EDGE) stage_tables_256(stab, A.b.tab, g.npml);  // the only workgroup barrier
    const int e = 4 * lb + w;
    if (e >= nst) return;
    const int nsz = (g.pnz + FR - 1) / FR, l = strips[e];
    const int tx = l % g.ntx, rest = l / g.ntx, sz = rest % nsz, s = rest / nsz;
    const FwiShotDev sh = shots[s];
    if (k >= sh.nt) return;
    if constexpr (EDGE) {  // the body compiled for the strip's rules, as in k_fwd_tile
        const int c = fwi_adj_strip_class(g, sh, tx * TX, sz * FR);
        if ((c & ~FWI_CLS_X) == 0)
            adj_tile_rows<EDGE, false, FWI_CLS_X>(A, sh, stab, Wsh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else if ((c & ~FWI_CLS_Z) == 0)
            adj_tile_rows<EDGE, false, FWI_CLS_Z>(A, sh, stab, Wsh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
        else
            adj_tile_rows<EDGE>(A, sh, stab, Wsh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    } else {
        adj_tile_rows<EDGE>(A, sh, stab, Wsh[w], Ush[w], s, k, tx * TX, sz * FR, lane);
    }
}

// The adjoint strips beside the fused backward step's interior launch.  A body
// compiled for every rule takes the most registers (86 VGPRs with 4-row
// strips, 116 with 8-row ones), so the strips whose rules are all on one side
// run lighter bodies (63 / 83 VGPRs): LIGHT covers strips whose rules are all
// on the x side (interior strips included: every x rule then resolves to the
// compact form) or all on the z side with the receiver rows; the others, the
// corner strips, run the body with every rule.  The host puts each strip in
// its list with fwi_adj_strip_class.
// One adjoint strip l of those lists at step k (CPML tables in stab)
template <bool LIGHT>
__device__ __forceinline__ void adj_split_strip(const FwiAdjArgs &A, const FwiShotDev *__restrict__ shots, int l,
                                                int k, int lane, float (*Ws)[HC], float (*Us)[HC], const float *stab)
{
    const FwiGeomDev &g = A.g;
    const int nsz = (g.pnz + FR - 1) / FR;
    const int tx = l % g.ntx, rest = l / g.ntx, sz = rest % nsz, s = rest / nsz;
    const FwiShotDev sh = shots[s];
    if (k >= sh.nt) return;
    if constexpr (LIGHT) {
        const int c = fwi_adj_strip_class(g, sh, tx * TX, sz * FR);
        if ((c & ~FWI_CLS_X) == 0)
            adj_tile_rows<true, false, FWI_CLS_X>(A, sh, stab, Ws, Us, s, k, tx * TX, sz * FR, lane);
        else if ((c & ~FWI_CLS_Z) == 0)
            adj_tile_rows<true, false, FWI_CLS_Z>(A, sh, stab, Ws, Us, s, k, tx * TX, sz * FR, lane);
        else
            adj_tile_rows<true, false, FWI_CLS_Z | FWI_CLS_RS>(A, sh, stab, Ws, Us, s, k, tx * TX, sz * FR, lane);
    } else {
        adj_tile_rows<true>(A, sh, stab, Ws, Us, s, k, tx * TX, sz * FR, lane);
    }
}

// Every adjoint strip of the other tiles in one launch, so that the
// reconstruction after it follows on the same stream: the corner strips (lh,
// the longer waves) in the first blocks, dealt round the XCDs, then the light
// strips (ll) in XCD-contiguous order.  Each part's block count is a multiple
// of 8, so the light part's blocks keep blockIdx & 7 as their XCD.
__global__ __launch_bounds__(256) void k_adj_tile_gen(const FwiAdjArgs *__restrict__ Ap,
                                                      const FwiShotDev *__restrict__ shots,
                                                      const int *__restrict__ lh, int nh,
                                                      const int *__restrict__ ll, int nl, const int *kbase, int koff)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiAdjArgs &A = *Ap;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int bh = fwi_gen_blocks(nh), b = blockIdx.x;
    __shared__ float Wsh[4][FR][HC];
    __shared__ float Ush[4][FR][HC];
    __shared__ float stab[FWI_TABN];
    if (b < bh) {
        if (4 * b >= nh) return;  // uniform: before the barrier
        stage_tables_256(stab, A.b.tab, A.g.npml);
        const int e = 4 * b + w;
        if (e < nh) adj_split_strip<false>(A, shots, lh[e], k, lane, Wsh[w], Ush[w], stab);
    } else {
        const int lb = xcd_remap(b - bh, (nl + 3) >> 2);
        if (lb < 0) return;
        stage_tables_256(stab, A.b.tab, A.g.npml);
        const int e = 4 * lb + w;
        if (e < nl) adj_split_strip<true>(A, shots, ll[e], k, lane, Wsh[w], Ush[w], stab);
    }
}

// ============================================================================
// reconstruction (physical grid) + zero-lag imaging
// ============================================================================
__global__ __launch_bounds__(FWI_TX * FWI_MAXSG) void k_rec_step(FwiRecArgs A, const int *kbase, int koff)
{
    // step index: koff alone, or a device counter plus koff inside a captured graph
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const FwiGeomDev &g = A.g;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ntiles = g.bntx * g.bntz;
    const int lb = xcd_remap(blockIdx.x, ntiles * A.b.ngroups);
    if (lb < 0) return;
    const int grp = lb / ntiles, t = lb - grp * ntiles;
    const int tzi = t / g.bntx, txi = t - tzi * g.bntx;
    const int s = grp * A.b.sg + wv;
    // every wave of the workgroup runs the whole loop (the imaging sum over
    // the group's shots meets at a barrier per row); inactive waves add 0
    bool active = s < A.b.nshots;
    FwiShotDev sh;
    sh.nt = 0;
    if (active) {
        sh = A.b.shots[s];
        active = k < sh.nt;
    }

    __shared__ float lds[FWI_MAXSG][2][RB][72];
    __shared__ double red[FWI_MAXSG][TX];
    __shared__ double redb[2][RB][FWI_MAXSG][TX];  // one barrier per interior block
    float(*Ps)[72] = lds[wv][0];
    float(*Us)[72] = lds[wv][1];

    const int npml = g.npml, nx = g.nx, nz = g.nz, ld = g.ld, sg = A.b.sg;
    const int bx0 = txi * TX, bx = bx0 + lane, fx = bx + npml;
    const int z0 = tzi * g.tz, z1 = min(z0 + g.tz, nz);
    const bool colok = bx < nx;
    const bool live = active && colok;
    const long long so = (long long)(active ? s : 0) * A.b.fs;
    const int cur = k & 1;
    const float *pn = (cur ? A.P1 : A.P0) + so;  // p^k
    float *pp = (cur ? A.P0 : A.P1) + so;        // p^{k+1} -> p^{k-1}
    const float *mu = (cur ? A.M0 : A.M1) + so;  // mu^{k+1}
    const float *ekp = A.b.ekp;
    const float c0x = g.c0x, c1x = g.c1x, c0z = g.c0z, c1z = g.c1z;
    FWI_GLOBAL double *gacc = gbl(A.gacc) + (long long)grp * nx * nz;
    const bool img_on = fwi_image_step(A.img_jump, A.img_k0, A.img_k1, k);  // uniform per launch
    const float fsrc = active ? gbl(A.b.srcf)[sh.src_off + k] : 0.f;
    const int hq = halo_col(bx0, lane);

    float p0w = 0.f, p1w = 0.f, p2w = 0.f, p3w = 0.f;
    float v0w = 0.f, v1w = 0.f, v2w = 0.f, v3w = 0.f;
    float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;

    // iteration r needs p^k[r] (own column and halo); its output row i = r - 3
    // needs p^{k+1}[i], mu^{k+1}[i], ekp[i] and (wave 0) the FP64 accumulator
    const int rend = z1 + 2;
    auto fetch = [&](int r, float &fs, float &fh, float &fp, float &fm, float &fk, double &fg) {
        fs = fh = fp = fm = fk = 0.f;
        fg = 0.0;
        if (r > rend) return;
        if (live && r >= 0 && r < nz) fs = pn[(long long)r * ld + fx];
        if (active && r >= z0 && r < z1 && r >= 3 && r < nz - 3 && hq >= 0 && hq < nx)
            fh = pn[(long long)r * ld + npml + hq];
        const int io = r - 3;
        if (io >= z0 && io < z1) {
            const long long o = (long long)io * ld + fx;
            if (live) {
                fp = pp[o];
                fm = mu[o];
                fk = ekp[o];
            }
            if (wv == 0 && colok && img_on) fg = gacc[(long long)io * nx + bx];
        }
    };
    // x chain of the interior operator for the block's rows in [z0, z1) n [3, nz-3)
    auto xblock = [&](auto Nc, int rb, const float *bs, const float *bh, float *bu) {
        constexpr int N = decltype(Nc)::value;
        wave_sync();
#pragma unroll
        for (int d = 0; d < N; ++d) {
            const int r = rb + d;
            if (r >= z0 && r < z1 && r >= 3 && r < nz - 3) {
                Ps[d][lane + 3] = bs[d];
                if (lane < 3) Ps[d][lane] = bh[d];
                else if (lane < 6) Ps[d][lane + 64] = bh[d];
            }
        }
        wave_sync();
#pragma unroll
        for (int d = 0; d < N; ++d) {
            const int r = rb + d;
            if (r >= z0 && r < z1 && r >= 3 && r < nz - 3) {
                const float *P = Ps[d];
                Us[d][lane] = c0x * (P[lane + 2] - P[lane + 1]) + c1x * (P[lane + 3] - P[lane]);
                if (lane < 3)
                    Us[d][64 + lane] = c0x * (P[lane + 66] - P[lane + 65]) + c1x * (P[lane + 67] - P[lane + 64]);
            }
        }
        wave_sync();
#pragma unroll
        for (int d = 0; d < N; ++d) {
            const int r = rb + d;
            bu[d] = 0.f;
            if (r >= z0 && r < z1 && r >= 3 && r < nz - 3) {
                const float *U = Us[d];
                bu[d] = c0x * (U[lane + 2] - U[lane + 1]) + c1x * (U[lane + 3] - U[lane]);
            }
        }
    };
    auto body = [&](int r, float pr, float uxr, float pp_i, float mu_i, float ekp_i, double g_i) {
        p0w = p1w; p1w = p2w; p2w = p3w; p3w = pr;
        q0 = q1; q1 = q2; q2 = q3; q3 = uxr;
        const float vj = c0z * (p2w - p1w) + c1z * (p3w - p0w);  // v[r-1]; used only where all taps are inside
        v0w = v1w; v1w = v2w; v2w = v3w; v3w = vj;

        const int i = r - 3;
        if (i >= z0 && i < z1) {
            double img = 0.0;
            if (live) {
                const long long so2 = sh.strip_off + (long long)k * g.sl;
                const bool inner = (i >= 3 && i < nz - 3 && bx >= 3 && bx < nx - 3);
                float X;
                if (inner) {
                    float ux = q0;
                    if (i == sh.isz && fx == sh.isx) ux = ux + fsrc;
                    const float vz = c0z * (v2w - v1w) + c1z * (v3w - v0w);
                    X = ux + vz;
                } else {
                    X = gbl(A.xstrip)[so2 + frame_pos(i, bx, nx, nz)];
                }
                img = (double)mu_i * (double)X;
                if (k >= 1) {
                    float pm;
                    if (inner) pm = leapfrog(p0w, pp_i, ekp_i * X);
                    else pm = (k >= 2) ? gbl(A.pstrip)[sh.strip_off + (long long)(k - 2) * g.sl + frame_pos(i, bx, nx, nz)]
                                       : 0.f;
                    pp[(long long)i * ld + fx] = pm;
                }
            }
            red[wv][lane] = img;
            __syncthreads();
            if (wv == 0 && colok && img_on) {
                double acc = g_i;
                for (int w = 0; w < sg; ++w) acc += red[w][lane];
                gacc[(long long)i * nx + bx] = acc;
            }
            __syncthreads();
        }
    };
    auto gen_block = [&](auto Nc, int rb) {
        constexpr int N = decltype(Nc)::value;
        float bs[N], bh[N], bp[N], bm[N], bk[N], bu[N];
        double bg[N];
#pragma unroll
        for (int d = 0; d < N; ++d) fetch(rb + d, bs[d], bh[d], bp[d], bm[d], bk[d], bg[d]);
        if (active && rb + N - 1 >= z0 && rb < z1) xblock(Nc, rb, bs, bh, bu);
        else {
#pragma unroll
            for (int d = 0; d < N; ++d) bu[d] = 0.f;
        }
#pragma unroll
        for (int d = 0; d < N; ++d)
            if (rb + d <= rend) body(rb + d, bs[d], bu[d], bp[d], bm[d], bk[d], bg[d]);
    };
    // interior blocks: the decision depends on the tile and rows only (the
    // same for every wave of the workgroup, which meets at one barrier)
    const bool fastb = bx0 >= 3 && bx0 + 63 < nx - 3;
    int blk = 0;
    auto run_block = [&](int rb, const float *bs, const float *bh, const float *bp, const float *bm,
                         const float *bk, const double *bg) {
        const int ilo = rb - 3, ihi = rb + RB - 4;
        if (fastb && ilo >= 3 && ihi < nz - 3) {
            float bu[RB];
            if (active) xblock(std::integral_constant<int, RB>{}, rb, bs, bh, bu);
            double(*rd)[FWI_MAXSG][TX] = redb[blk & 1];
#pragma unroll
            for (int d = 0; d < RB; ++d) {
                p0w = p1w; p1w = p2w; p2w = p3w; p3w = bs[d];
                q0 = q1; q1 = q2; q2 = q3; q3 = active ? bu[d] : 0.f;
                const float vj = c0z * (p2w - p1w) + c1z * (p3w - p0w);
                v0w = v1w; v1w = v2w; v2w = v3w; v3w = vj;
                const int i = rb + d - 3;
                double img = 0.0;
                if (active) {
                    float ux = q0;
                    if (i == sh.isz && fx == sh.isx) ux = ux + fsrc;
                    const float vz = c0z * (v2w - v1w) + c1z * (v3w - v0w);
                    const float X = ux + vz;
                    img = (double)bm[d] * (double)X;
                    if (k >= 1) pp[i * ld + fx] = leapfrog(p0w, bp[d], bk[d] * X);
                }
                rd[d][wv][lane] = img;
            }
            __syncthreads();
            if (wv == 0 && img_on) {
#pragma unroll
                for (int d = 0; d < RB; ++d) {
                    double acc = bg[d];
                    for (int w = 0; w < sg; ++w) acc += rd[d][w][lane];
                    gacc[(rb + d - 3) * nx + bx] = acc;
                }
            }
            ++blk;
        } else {
            float bu[RB];
            if (active) xblock(std::integral_constant<int, RB>{}, rb, bs, bh, bu);
            else {
#pragma unroll
                for (int d = 0; d < RB; ++d) bu[d] = 0.f;
            }
#pragma unroll
            for (int d = 0; d < RB; ++d) body(rb + d, bs[d], bu[d], bp[d], bm[d], bk[d], bg[d]);
        }
    };
    using I3 = std::integral_constant<int, 3>;
    gen_block(I3{}, z0 - 3);
    gen_block(I3{}, z0);
    int rb = z0 + 3;
    const int rmain = z1 - 1;
    if (rb + RB - 1 <= rmain) {
        float As[RB], Ah[RB], Ap[RB], Am[RB], Ak[RB], Bs[RB], Bh[RB], Bp[RB], Bm[RB], Bk[RB];
        double Ag[RB], Bg[RB];
#pragma unroll
        for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ap[d], Am[d], Ak[d], Ag[d]);
        for (;;) {
            const bool nb = rb + 2 * RB - 1 <= rmain;
            if (nb && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, Bs[d], Bh[d], Bp[d], Bm[d], Bk[d], Bg[d]);
            }
            run_block(rb, As, Ah, Ap, Am, Ak, Ag);
            rb += RB;
            if (!nb) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, Bs[d], Bh[d], Bp[d], Bm[d], Bk[d], Bg[d]);
            }
            const bool na = rb + 2 * RB - 1 <= rmain;
            if (na && FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + RB + d, As[d], Ah[d], Ap[d], Am[d], Ak[d], Ag[d]);
            }
            run_block(rb, Bs, Bh, Bp, Bm, Bk, Bg);
            rb += RB;
            if (!na) break;
            if (!FWI_DOUBLEBUF) {
#pragma unroll
                for (int d = 0; d < RB; ++d) fetch(rb + d, As[d], Ah[d], Ap[d], Am[d], Ak[d], Ag[d]);
            }
        }
    }
    for (; rb <= rend; rb += 3) gen_block(I3{}, rb);
    __syncthreads();  // the last interior block's partial sums are read before the waves leave
}

// ============================================================================
// reconstruction + zero-lag imaging, 2-D tile form on the physical grid: one
// workgroup (4 waves) per 64 x FR physical tile and group of shots; wave w
// takes the shots s_lo + w, s_lo + w + 4, ...  Per shot a wave reconstructs
// p^{k-1} on the tile's rows (interior operator with the z chain in registers
// and the x chain through wave-private LDS; frame cells from the saved
// strips) and adds mu^{k+1} X^k to FP64 accumulators held in registers.  At
// the end the four partial images meet in LDS (fixed order, so the sum is
// deterministic) and the group's image is added to its FP64 accumulator in
// HBM once per step.  Same arithmetic as k_rec_step.
// ============================================================================
#define RR FWI_REC_ROWS
#define RPR (RR + 6)
#define RVN (RR + 3)
#define RHP ((6 * RR + 63) / 64)
static_assert(RR % 4 == 0 && 3 * RR <= 64, "reconstruction rows per wave");

// The loads of one shot's reconstruction step on a tile: p^k (rows with a
// 3-row halo, the x halo), p^{k+1} and, unless FUSED, mu^{k+1}.  They do not
// depend on the adjoint step, so the fused backward step issues them before
// the adjoint's arithmetic.  bx0 may be negative (tiles of the padded grid
// start npml columns left of the grid).
// LATE: the tile images step k+1 at step k, mu^{k+2} X^{k+1} (mu^{k+2} is
// read-only during step k; X^{k+1} was kept in xg by step k+1), so that it
// does not wait for the adjoint step k of its cells; rec_tile_shot loads
// those two (late_load), and nothing is loaded here at k = -1 (the image of
// step 0, after the last step).
struct RecIn {
    float p[RPR], hv[RHP], pq[RR], pm[RR];
};
template <bool EDGE, bool FUSED = false, bool LATE = false>
__device__ __forceinline__ void rec_tile_load(const FwiRecArgs &A, int s, int k, int bx0, int r0, int lane,
                                              RecIn &in)
{
    if (LATE && k < 0) return;
    const FwiGeomDev &g = A.g;
    const int npml = g.npml, nx = g.nx, nz = g.nz, ld = g.ld;
    const int bx = bx0 + lane, fx = bx + npml;
    const bool colok = !EDGE || (bx >= 0 && bx < nx);
    const long long so = (long long)s * A.b.fs;
    const int cur = k & 1;
    const float *pn = (cur ? A.P1 : A.P0) + so;  // p^k
    const float *pp = (cur ? A.P0 : A.P1) + so;  // p^{k+1}
    const float *mu = (cur ? A.M0 : A.M1) + so;  // mu^{k+1}
    const unsigned fbytes = (unsigned)(A.b.fs * sizeof(float));
    const brsrc rn = buf_rsrc(pn, fbytes), rq = buf_rsrc(pp, fbytes), rm = buf_rsrc(mu, fbytes);
    const int vo = (colok ? fx : npml) * 4, rowb = ld * 4;
#pragma unroll
    for (int j = 0; j < RPR; ++j) {
        const int zr = r0 - 3 + j;
        in.p[j] = EDGE ? buf_ld_row_sel(rn, vo, zr, nz, rowb) : buf_ld(rn, vo, zr * rowb);
        if (EDGE && !colok) in.p[j] = 0.f;
    }
    XHalo<RHP> xh;
    xhalo_plan(xh, lane, bx0, RR);
#pragma unroll
    for (int q = 0; q < RHP; ++q) {
        bool ok = xh.in[q];
        if (EDGE) ok = ok && r0 + xh.d[q] < nz && xh.col[q] >= 0 && xh.col[q] < nx;
        in.hv[q] = buf_ld(rn, ok ? (xh.d[q] * ld + xh.col[q] + npml) * 4 : vo, r0 * rowb);
        in.hv[q] = ok ? in.hv[q] : 0.f;
    }
#pragma unroll
    for (int d = 0; d < RR; ++d) {
        const int i = r0 + d;
        in.pq[d] = EDGE ? buf_ld_row_sel(rq, vo, i, nz, rowb) : buf_ld(rq, vo, i * rowb);
        if (!FUSED && !LATE) in.pm[d] = EDGE ? buf_ld_row_sel(rm, vo, i, nz, rowb) : buf_ld(rm, vo, i * rowb);
    }
}

// mu^{k+2} and X^{k+1} of a tile's rows (LATE imaging)
template <bool EDGE>
__device__ __forceinline__ void late_load(const FwiRecArgs &A, int s, int k, int bx0, int r0, int lane, float *mu2,
                                          float *xq)
{
    const FwiGeomDev &g = A.g;
    const int npml = g.npml, nx = g.nx, nz = g.nz, ld = g.ld;
    const int bx = bx0 + lane;
    const bool colok = !EDGE || (bx >= 0 && bx < nx);
    const float *mu = ((k & 1) ? A.M1 : A.M0) + (long long)s * A.b.fs;  // mu^{k+2}
    const brsrc rm = buf_rsrc(mu, (unsigned)(A.b.fs * sizeof(float)));
    const brsrc rx =
        buf_rsrc(A.xg + ((k + 1) & 1) * A.xg_b + (long long)s * nx * nz, (unsigned)(sizeof(float) * nx * nz));
    const int vo = (colok ? bx + npml : npml) * 4, rowb = ld * 4, xo = (colok ? bx : 0) * 4;
#pragma unroll
    for (int d = 0; d < RR; ++d) {
        const int i = r0 + d;
        mu2[d] = EDGE ? buf_ld_row_sel(rm, vo, i, nz, rowb) : buf_ld(rm, vo, i * rowb);
        xq[d] = buf_ld_row_sel(rx, xo, i, nz, nx * 4);
        xq[d] = colok ? xq[d] : 0.f;
    }
}

// One shot's reconstruction + imaging step on a tile from the loads of
// rec_tile_load.  FUSED: mu^{k+1} of the rows comes from the caller (mu1).
// LATE: the image term is that of step k+1 (when the shot has one), and X^k
// goes to xg for step k-1.
template <bool EDGE, bool FUSED = false, bool LATE = false>
__device__ __forceinline__ void rec_tile_shot(const FwiRecArgs &A, const FwiShotDev &sh, int s, int k, int bx0,
                                              int r0, int lane, float (*Ps)[HC], float (*Us)[HC], const float *pk,
                                              double *acc, const RecIn &in, const float *mu1 = nullptr)
{
    const FwiGeomDev &g = A.g;
    const int npml = g.npml, nx = g.nx, nz = g.nz, ld = g.ld;
    const int bx = bx0 + lane, fx = bx + npml;
    const bool colok = !EDGE || (bx >= 0 && bx < nx);
    // LATE: the term of step k+1, in the order the other tiles add theirs
    const bool late_term = LATE && k + 1 < sh.nt;
    float mu2[RR], xq[RR];
    if (LATE && k < 0) {
        late_load<EDGE>(A, s, k, bx0, r0, lane, mu2, xq);
#pragma unroll
        for (int d = 0; d < RR; ++d)
            if (colok && r0 + d < nz) acc[d] += (double)mu2[d] * (double)xq[d];
        return;
    }
    const long long so = (long long)s * A.b.fs;
    const int cur = k & 1;
    float *pp = (cur ? A.P0 : A.P1) + so;        // p^{k+1} -> p^{k-1}
    const unsigned fbytes = (unsigned)(A.b.fs * sizeof(float));
    const brsrc rq = buf_rsrc(pp, fbytes);
    const int vo = (colok ? fx : npml) * 4, rowb = ld * 4;
    const float *p = in.p, *pq = in.pq;
    const float *pm = FUSED ? mu1 : in.pm;
    XHalo<RHP> xh;
    xhalo_plan(xh, lane, bx0, RR);
    const float *hv = in.hv;

    wave_sync();  // the previous shot's reads of Ps / Us are done
#pragma unroll
    for (int d = 0; d < RR; ++d) Ps[d][lane + 3] = p[d + 3];
#pragma unroll
    for (int q = 0; q < RHP; ++q) Ps[xh.d[q]][xhalo_lds_col(xh, q, bx0)] = hv[q];
    wave_sync();
    const float c0x = g.c0x, c1x = g.c1x, c0z = g.c0z, c1z = g.c1z;
#pragma unroll
    for (int d = 0; d < RR; ++d) {
        const float *P = Ps[d];
        Us[d][lane] = c0x * (P[lane + 2] - P[lane + 1]) + c1x * (P[lane + 3] - P[lane]);
    }
    {
        const int ud0 = lane / 3;
        const bool uok = lane < 3 * RR;
        const int ud = uok ? ud0 : 0, uc = uok ? 64 + lane - 3 * ud0 : 64;
        const float *P = Ps[ud];
        Us[ud][uok ? uc : HJ] = c0x * (P[uc + 2] - P[uc + 1]) + c1x * (P[uc + 3] - P[uc]);
    }
    wave_sync();
    float v[RVN];
#pragma unroll
    for (int n = 0; n < RVN; ++n) v[n] = c0z * (p[n + 2] - p[n + 1]) + c1z * (p[n + 3] - p[n]);

    const long long so2 = sh.strip_off + (long long)k * g.sl;
    // general tiles: the frame values and the source sample are read before the
    // first store (vector memory operations retire in order)
    float xsv[RR], psv[RR], fsrc = 0.f;
    if (EDGE) {
        const long long so0 = sh.strip_off + (long long)(k - 2) * g.sl;  // p^{k-1} was saved at step k-2
#pragma unroll
        for (int d = 0; d < RR; ++d) {
            const int i = r0 + d;
            const bool fr = colok && i < nz && !(i >= 3 && i < nz - 3 && bx >= 3 && bx < nx - 3);
            const int pos = fr ? frame_pos(i, bx, nx, nz) : 0;
            xsv[d] = gbl(A.xstrip)[so2 + pos];
            psv[d] = ld_sel(A.pstrip, so0 + pos, k >= 2, 0.f);
        }
        // used only on the source row
        fsrc = gbl(A.b.srcf)[sh.src_off + ((sh.isz >= r0 && sh.isz < r0 + RR) ? k : 0)];
    }
    if (LATE) late_load<EDGE>(A, s, k, bx0, r0, lane, mu2, xq);
    const brsrc rxk =
        buf_rsrc(A.xg + (k & 1) * A.xg_b + (long long)s * nx * nz, (unsigned)(sizeof(float) * nx * nz));
#pragma unroll
    for (int d = 0; d < RR; ++d) {
        const int i = r0 + d;
        if (EDGE && i >= nz) break;
        const float *U = Us[d];
        float ux = c0x * (U[lane + 2] - U[lane + 1]) + c1x * (U[lane + 3] - U[lane]);
        const float vz = c0z * (v[d + 2] - v[d + 1]) + c1z * (v[d + 3] - v[d]);
        if (EDGE && i == sh.isz) ux = (fx == sh.isx) ? ux + fsrc : ux;
        float X = ux + vz;
        bool inner = true;
        if (EDGE) {
            inner = i >= 3 && i < nz - 3 && bx >= 3 && bx < nx - 3;
            if (!inner && colok) X = xsv[d];
        }
        if (LATE) {
            if (colok && late_term) acc[d] += (double)mu2[d] * (double)xq[d];
            if (colok) buf_st(X, rxk, bx * 4, i * nx * 4);
        } else if (colok) {
            acc[d] += (double)pm[d] * (double)X;
        }
        if (k >= 1) {
            float pnew;
            if (!EDGE || inner) pnew = leapfrog(p[d + 3], pq[d], pk[d] * X);
            else pnew = colok ? psv[d] : 0.f;
            if (colok) buf_st(pnew, rq, vo, i * rowb);
        }
    }
}


// items (group, tile) of the lists built with fwi_rec_tile_interior (an item
// is interior when the tile is interior for every shot of the group).
// PADDED: the tiles are those of the padded grid (64 x FR from column 0, as
// the fused backward step's), so bx0 = the tile's first column - npml.
// LDS of a reconstruction item: per wave, staged p rows and u rows while
// propagating, then its partial image (FP64)
typedef float RecLds[4][2][RR][HC];
static_assert(sizeof(float[2][RR][HC]) >= RR * TX * sizeof(double), "partial image fits the wave's rows");

// One reconstruction item l = (group, tile) at step k (LATE: see rec_tile_load)
template <bool EDGE, bool PADDED, bool LATE>
__device__ __forceinline__ void rec_item(const FwiRecArgs &A, const FwiShotDev *__restrict__ shots, int l, int k,
                                         int gsize, int lane, int w, RecLds &lds)
{
    const FwiGeomDev &g = A.g;
    const int ncol = PADDED ? g.ntx : g.bntx;
    const int ntz = PADDED ? (g.pnz + RR - 1) / RR : (g.nz + RR - 1) / RR;
    const int ntiles = ncol * ntz;
    const int grp = l / ntiles, t = l - grp * ntiles;
    const int tzi = t / ncol, txi = t - tzi * ncol;
    const int bx0 = txi * TX - (PADDED ? g.npml : 0), r0 = tzi * RR;
    const int nx = g.nx, nz = g.nz, ld = g.ld, npml = g.npml;
    const int bx = bx0 + lane;
    const bool colok = bx >= 0 && bx < nx;
    // ekp of the tile's rows, shared by every shot
    float pk[RR];
    {
        const brsrc rk = buf_rsrc(A.b.ekp, (unsigned)(A.b.fs * sizeof(float)));
        const int vo = (colok ? bx + npml : npml) * 4;
#pragma unroll
        for (int d = 0; d < RR; ++d) pk[d] = buf_ld_row_sel(rk, vo, r0 + d, nz, ld * 4);
    }
    double acc[RR];
#pragma unroll
    for (int d = 0; d < RR; ++d) acc[d] = 0.0;
    const int s_lo = grp * gsize, s_hi = min(s_lo + gsize, A.b.nshots);
    for (int s = s_lo + w; s < s_hi; s += 4) {
        const FwiShotDev sh = shots[s];
        if (k >= sh.nt) continue;
        RecIn rin;
        rec_tile_load<EDGE, false, LATE>(A, s, k, bx0, r0, lane, rin);
        rec_tile_shot<EDGE, false, LATE>(A, sh, s, k, bx0, r0, lane, lds[w][0], lds[w][1], pk, acc, rin);
    }
    // steps outside the imaging window only reconstruct (uniform per launch)
    if (!fwi_image_step(A.img_jump, A.img_k0, A.img_k1, LATE ? k + 1 : k)) return;
    // partial images of the four waves -> LDS; wave w then finishes rows 2w, 2w+1
    wave_sync();
    double(*red)[TX] = reinterpret_cast<double(*)[TX]>(&lds[w][0][0][0]);
#pragma unroll
    for (int d = 0; d < RR; ++d) red[d][lane] = acc[d];
    __syncthreads();
    FWI_GLOBAL double *gacc = gbl(A.gacc) + (long long)grp * nx * nz;
#pragma unroll
    for (int e = 0; e < RR / 4; ++e) {
        const int d = w * (RR / 4) + e, i = r0 + d;
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) sum += reinterpret_cast<double(*)[TX]>(&lds[q][0][0][0])[d][lane];
        if (colok && i < nz) gacc[(long long)i * nx + bx] += sum;
    }
}

template <bool EDGE, bool PADDED = false>
__device__ __forceinline__ void rec_tile_kernel(const FwiRecArgs *__restrict__ Ap, const FwiShotDev *__restrict__ shots,
                                                const int *__restrict__ items, int nit, const int *kbase, int koff,
                                                int gsize)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, nit);
    if (lb < 0) return;
    __shared__ __attribute__((aligned(16))) RecLds lds;
    rec_item<EDGE, PADDED, false>(*Ap, shots, items[lb], k, gsize, lane, w, lds);
}

// Occupancy targets: the interior items fit 80 VGPRs without spilling when the
// compiler aims at 6 waves per SIMD (its default schedule takes 98, 4 waves);
// the general items fit 4 waves (123 VGPRs) without spilling.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_rec_tile_int(
    const FwiRecArgs *__restrict__ Ap, const FwiShotDev *__restrict__ shots, const int *__restrict__ items, int nit,
    const int *kbase, int koff, int gsize)
{
    rec_tile_kernel<false>(Ap, shots, items, nit, kbase, koff, gsize);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_rec_tile_gen(const FwiRecArgs *__restrict__ Ap,
                                                      const FwiShotDev *__restrict__ shots,
                                                      const int *__restrict__ items, int nit, const int *kbase,
                                                      int koff, int gsize)
{
    rec_tile_kernel<true>(Ap, shots, items, nit, kbase, koff, gsize);
}
// general items (group, padded tile) beside the fused backward step's interior ones
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_rec_tile_pad(const FwiRecArgs *__restrict__ Ap,
                                                      const FwiShotDev *__restrict__ shots,
                                                      const int *__restrict__ items, int nit, const int *kbase,
                                                      int koff, int gsize)
{
    rec_tile_kernel<true, true>(Ap, shots, items, nit, kbase, koff, gsize);
}

// ============================================================================
// fused backward step on interior items: adjoint step k, then reconstruction
// + imaging step k, of a shot group on one 64 x FR tile of the padded grid
// that is interior for both steps and every shot of the group.  The imaging
// takes mu^{k+1} from the adjoint's registers and ekp is read once per tile
// for the whole group, instead of mu^{k+1} re-read and ekp read per shot by
// the separate kernels.  Each wave runs the shots w, w+4, ... of the group
// and the four partial images are summed in wave order, as in
// rec_tile_kernel, so every cell's image sum is the same bit for bit.  The
// other tiles run as k_adj_tile strips, then k_rec_tile_pad items (on such
// tiles the fused form ran slower than the two kernels: 2 waves per SIMD).
// Items: group-major (group, row strip, tile column) of the padded grid.
// ============================================================================
// Occupancy: 123 VGPRs without spilling at 4 waves per SIMD.  A cap of 112
// (so that a 160-VGPR general adjoint wave fits beside three of these waves)
// spills 8 VGPRs and costs more than the co-residency gains.
// One fused item l = (group, tile) at step k; lds per wave: the adjoint's
// w' / u~ rows, then the reconstruction's p / u rows, then its partial image
// (FP64)
__device__ __forceinline__ void bwd_int_item(const FwiAdjArgs &A, const FwiRecArgs &R,
                                             const FwiShotDev *__restrict__ shots, int l, int k, int gsize, int lane,
                                             int w, RecLds &lds)
{
    static_assert(FR == RR, "the adjoint and reconstruction rows of a tile coincide");
    const FwiGeomDev &g = A.g;
    const int nsz = (g.pnz + FR - 1) / FR, ntiles = g.ntx * nsz;
    const int grp = l / ntiles, t = l - grp * ntiles;
    const int sz = t / g.ntx, tx = t - sz * g.ntx;
    const int x0 = tx * TX, r0 = sz * FR;
    const int npml = g.npml, nx = g.nx, ld = g.ld;
    const int x = x0 + lane, bx = x - npml;
    // ekp of the tile's rows, shared by every shot and both steps
    float ek[FR];
    {
        const brsrc rk = buf_rsrc(A.b.ekp, (unsigned)(A.b.fs * sizeof(float)));
#pragma unroll
        for (int d = 0; d < FR; ++d) ek[d] = buf_ld(rk, x * 4, (r0 + d) * ld * 4);
    }
    double acc[FR];
#pragma unroll
    for (int d = 0; d < FR; ++d) acc[d] = 0.0;
    const int s_lo = grp * gsize, s_hi = min(s_lo + gsize, A.b.nshots);
    for (int s = s_lo + w; s < s_hi; s += 4) {
        const FwiShotDev sh = shots[s];
        if (k >= sh.nt) continue;
        // the reconstruction's loads first: they do not depend on the adjoint
        // step, so they are in flight during its arithmetic
        RecIn rin;
        rec_tile_load<false, true>(R, s, k, x0 - npml, r0, lane, rin);
        float mu1[FR];
        wave_sync();  // the previous shot's reads of the wave's LDS rows are done
        adj_tile_rows<false, true>(A, sh, nullptr, lds[w][0], lds[w][1], s, k, x0, r0, lane, ek, mu1);
        rec_tile_shot<false, true>(R, sh, s, k, x0 - npml, r0, lane, lds[w][0], lds[w][1], ek, acc, rin, mu1);
    }
    // steps outside the imaging window only propagate (uniform per launch)
    if (!fwi_image_step(R.img_jump, R.img_k0, R.img_k1, k)) return;
    wave_sync();
    double(*red)[TX] = reinterpret_cast<double(*)[TX]>(&lds[w][0][0][0]);
#pragma unroll
    for (int d = 0; d < FR; ++d) red[d][lane] = acc[d];
    __syncthreads();
    FWI_GLOBAL double *gacc = gbl(R.gacc) + (long long)grp * nx * g.nz;
#pragma unroll
    for (int e = 0; e < FR / 4; ++e) {
        const int d = w * (FR / 4) + e, i = r0 + d;
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) sum += reinterpret_cast<double(*)[TX]>(&lds[q][0][0][0])[d][lane];
        gacc[(long long)i * nx + bx] += sum;
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bwd_tile_int(
    const FwiAdjArgs *__restrict__ Aa, const FwiRecArgs *__restrict__ Ar, const FwiShotDev *__restrict__ shots,
    const int *__restrict__ items, int nit, const int *kbase, int koff, int gsize)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < 0) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, nit);
    if (lb < 0) return;
    __shared__ __attribute__((aligned(16))) RecLds lds;
    bwd_int_item(*Aa, *Ar, shots, items[lb], k, gsize, lane, w, lds);
}

// The fused backward step as one launch, so that no step waits on another
// stream: role 0 the interior items (l0, one per workgroup), 1 the other
// tiles' x-side and z-side adjoint strips (l1, four per workgroup), 2 their
// corner strips (l2), 3 their reconstruction items (l3), which image one step
// late (rec_tile_load LATE) and so read nothing the same launch writes.  At
// k = -1 only role 3 works: the image of step 0.  work[b] = (role << 28) |
// workgroup index in the role's list, -1 idle, as in k_fwd_mixed.  The roles
// share one LDS block; a workgroup runs one role, so the workgroup barriers
// inside each are reached by all its waves.
#define FWI_LDS_MIXED (2 * 4 * FR * HC + FWI_TABN)
static_assert(sizeof(RecLds) <= FWI_LDS_MIXED * sizeof(float), "every role fits the shared block");
__global__ __launch_bounds__(256) void k_bwd_mixed(const FwiAdjArgs *__restrict__ Aa, const FwiRecArgs *__restrict__ Ar,
                                                   const FwiShotDev *__restrict__ shots, const int *__restrict__ work,
                                                   const int *__restrict__ l0, const int *__restrict__ l1, int n1,
                                                   const int *__restrict__ l2, int n2, const int *__restrict__ l3,
                                                   const int *kbase, int koff, int gsize)
{
    const int k = (kbase ? *kbase : 0) + koff;
    if (k < -1) return;
    const int wv = work[blockIdx.x];
    if (wv < 0) return;
    const int role = wv >> 28, lb = wv & 0x0FFFFFFF;
    if (k < 0 && role != 3) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __shared__ __attribute__((aligned(16))) float smem[FWI_LDS_MIXED];
    if (role == 0) {
        bwd_int_item(*Aa, *Ar, shots, l0[lb], k, gsize, lane, w, *reinterpret_cast<RecLds *>(smem));
    } else if (role == 3) {
        rec_item<true, true, true>(*Ar, shots, l3[lb], k, gsize, lane, w, *reinterpret_cast<RecLds *>(smem));
    } else {
        float(*Ws)[FR][HC] = reinterpret_cast<float(*)[FR][HC]>(smem);
        float(*Us)[FR][HC] = Ws + 4;
        float *stab = smem + 2 * 4 * FR * HC;
        stage_tables_256(stab, Aa->b.tab, Aa->g.npml);
        const int e = 4 * lb + w;
        if (role == 1) {
            if (e < n1) adj_split_strip<true>(*Aa, shots, l1[e], k, lane, Ws[w], Us[w], stab);
        } else {
            if (e < n2) adj_split_strip<false>(*Aa, shots, l2[e], k, lane, Ws[w], Us[w], stab);
        }
    }
}

// ============================================================================
// small kernels
// ============================================================================
// moves the step counter of a captured block of steps
__global__ void k_advance(int *kbase, int delta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) kbase[0] += delta;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

__global__ void k_fill_f32(float *p, long long n, float v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double *p, long long n, double v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// time-major [nt][nr] <-> trace-major [nr][nt] for one shot
__global__ void k_transpose(const float *in, float *out, int rows, int cols)
{
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int r = by + yy, c = bx + threadIdx.x;
        if (r < rows && c < cols) tile[yy][threadIdx.x] = in[(long long)r * cols + c];
    }
    __syncthreads();
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int r = bx + yy, c = by + threadIdx.x;  // out is cols x rows
        if (r < cols && c < rows) out[(long long)r * rows + c] = tile[threadIdx.x][yy];
    }
}

// res = syn - obs; partial sums of 1/2 res^2 in double, one per block
__global__ void k_misfit_l2(const float *syn, const float *obs, float *res, long long n, double *partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float rr = syn[i] - obs[i];
        res[i] = rr;
        acc += 0.5 * (double)rr * (double)rr;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// grad = 2 * sum_groups gacc / vel
__global__ void k_grad_finalize(const double *gacc, int ngroups, long long n, const float *vel, double *gk,
                                double *grad)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int gi = 0; gi < ngroups; ++gi) acc += gacc[(long long)gi * n + i];
        gk[i] = acc;
        grad[i] = 2.0 * acc / (double)vel[i];
    }
}

// ============================================================================
// per-step kernels with the reference's loop structure (unit parity)
// p[pnz][pnx], u[pnz][pnx+1], v[pnz+1][pnx], ux/vz[pnz][pnx], dense rows.
// ============================================================================

// phase = index of the reference loop in a_lap_opt_2d_pml (0..11, fwi_2d.c:775-862)
// or 20..23 for a_lap_opt_2d (fwi_2d.c:735-756)
__global__ void k_ref_phase(FwiRefArgs a, int phase)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
    const int pnx = a.pnx, pnz = a.pnz, npml = a.npml, T = 2 * npml + 1;
    const int unx = pnx + 1, unz = pnz, vnx = pnx, vnz = pnz + 1, n1 = npml + 1;
    const float *p = a.p, *tab = a.tab;
    float *u = a.u, *v = a.v, *ux = a.ux, *vz = a.vz;
#define P_(z, x) p[(z) * pnx + (x)]
#define U_(z, x) u[(z) * unx + (x)]
#define V_(z, x) v[(z) * vnx + (x)]
#define UX_(z, x) ux[(z) * pnx + (x)]
#define VZ_(z, x) vz[(z) * pnx + (x)]
    switch (phase) {
    case 0:
        if (iz < unz && ix >= 2 && ix < unx - 2)
            U_(iz, ix) = a.c0x * (P_(iz, ix) - P_(iz, ix - 1)) + a.c1x * (P_(iz, ix + 1) - P_(iz, ix - 2));
        break;
    case 1:
        if (iz < unz && ix == 0) {
            U_(iz, 1) = a.ix * (P_(iz, 1) - P_(iz, 0));
            U_(iz, unx - 2) = a.ix * (P_(iz, pnx - 1) - P_(iz, pnx - 2));
        }
        break;
    case 2:
        if (iz < unz && ix < npml + 1) {
            a.u0[iz * n1 + ix] = TB(1, 2 * ix) * a.u0[iz * n1 + ix] + TB(0, 2 * ix) * U_(iz, ix);
            U_(iz, ix) += a.u0[iz * n1 + ix];
            a.u1[iz * n1 + ix] = TB(3, 2 * ix) * a.u1[iz * n1 + ix] + TB(2, 2 * ix) * U_(iz, ix + unx - npml - 1);
            U_(iz, ix + unx - npml - 1) += a.u1[iz * n1 + ix];
        }
        break;
    case 3:
        if (iz >= 2 && iz < vnz - 2 && ix < vnx)
            V_(iz, ix) = a.c0z * (P_(iz, ix) - P_(iz - 1, ix)) + a.c1z * (P_(iz + 1, ix) - P_(iz - 2, ix));
        break;
    case 4:
        if (iz == 0 && ix < vnx) {
            V_(0, ix) = v_row0(P_(0, ix), P_(1, ix), P_(2, ix), P_(3, ix), a.iz840);
            V_(1, ix) = v_row1(P_(0, ix), P_(1, ix), P_(2, ix), P_(3, ix), a.iz840);
            V_(vnz - 2, ix) = (P_(pnz - 1, ix) - P_(pnz - 2, ix)) * a.iz;
        }
        break;
    case 5:
        if (iz < npml + 1 && ix < vnx) {
            a.v1[iz * vnx + ix] = TB(5, 2 * iz) * a.v1[iz * vnx + ix] + TB(4, 2 * iz) * V_(iz + vnz - npml - 1, ix);
            V_(iz + vnz - npml - 1, ix) += a.v1[iz * vnx + ix];
        }
        break;
    case 6:
        if (iz < pnz && ix >= 1 && ix < pnx - 1)
            UX_(iz, ix) = a.c0x * (U_(iz, ix + 1) - U_(iz, ix)) + a.c1x * (U_(iz, ix + 2) - U_(iz, ix - 1));
        break;
    case 7:
        if (iz < pnz && ix == 0) {
            UX_(iz, 0) = (U_(iz, 1) - U_(iz, 0)) * a.ix;
            UX_(iz, pnx - 1) = (U_(iz, unx - 1) - U_(iz, unx - 2)) * a.ix;
        }
        break;
    case 8:
        if (iz < pnz && ix < npml) {
            a.ux0[iz * npml + ix] = TB(1, 2 * ix + 1) * a.ux0[iz * npml + ix] + TB(0, 2 * ix + 1) * UX_(iz, ix);
            UX_(iz, ix) += a.ux0[iz * npml + ix];
            a.ux1[iz * npml + ix] = TB(3, 2 * ix + 1) * a.ux1[iz * npml + ix] + TB(2, 2 * ix + 1) * UX_(iz, ix + pnx - npml);
            UX_(iz, ix + pnx - npml) += a.ux1[iz * npml + ix];
        }
        break;
    case 9:
        if (iz == 0 && ix < pnx) {
            VZ_(0, ix) = vz_row0(V_(0, ix), V_(1, ix), V_(2, ix), V_(3, ix), V_(4, ix), a.iz24);
            VZ_(pnz - 1, ix) = (V_(vnz - 1, ix) - V_(vnz - 2, ix)) * a.iz;
        }
        break;
    case 10:
        if (iz >= 1 && iz < pnz - 1 && ix < pnx)
            VZ_(iz, ix) = a.c0z * (V_(iz + 1, ix) - V_(iz, ix)) + a.c1z * (V_(iz + 2, ix) - V_(iz - 1, ix));
        break;
    case 11:
        if (iz < npml && ix < pnx) {
            a.vz1[iz * pnx + ix] = TB(5, 2 * iz + 1) * a.vz1[iz * pnx + ix] + TB(4, 2 * iz + 1) * VZ_(iz + pnz - npml, ix);
            VZ_(iz + pnz - npml, ix) += a.vz1[iz * pnx + ix];
        }
        break;
    case 20:
        if (iz >= 2 && iz < unz - 2 && ix >= 2 && ix < unx - 2)
            U_(iz, ix) = a.c0x * (P_(iz, ix) - P_(iz, ix - 1)) + a.c1x * (P_(iz, ix + 1) - P_(iz, ix - 2));
        break;
    case 21:
        if (iz >= 2 && iz < vnz - 2 && ix >= 2 && ix < vnx - 2)
            V_(iz, ix) = a.c0z * (P_(iz, ix) - P_(iz - 1, ix)) + a.c1z * (P_(iz + 1, ix) - P_(iz - 2, ix));
        break;
    case 22:
        if (iz >= 3 && iz < pnz - 3 && ix >= 3 && ix < pnx - 3)
            UX_(iz, ix) = a.c0x * (U_(iz, ix + 1) - U_(iz, ix)) + a.c1x * (U_(iz, ix + 2) - U_(iz, ix - 1));
        break;
    case 23:
        if (iz >= 3 && iz < pnz - 3 && ix >= 3 && ix < pnx - 3)
            VZ_(iz, ix) = a.c0z * (V_(iz + 1, ix) - V_(iz, ix)) + a.c1z * (V_(iz + 2, ix) - V_(iz - 1, ix));
        break;
    default:
        break;
    }
#undef P_
#undef U_
#undef V_
#undef UX_
#undef VZ_
}

__global__ void k_ref_time_integral(float *pnew, const float *pnow, const float *pold, const float *pxx,
                                    const float *pzz, const float *ekp, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        pnew[i] = leapfrog(pnow[i], pold[i], ekp[i] * (pxx[i] + pzz[i]));
}

// write_bdr / read_bdr for one step (fwi_2d.c:891-910, 940-957)
// top: rows of CPML above the physical grid (0 with the free surface)
__global__ void k_ref_write_bdr(float *ud, float *lr, const float *pf, int nx, int nz, int npml, int top)
{
    const int pnx_f = nx + 2 * npml;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * nx) {
        const int i = t / nx, j = t - i * nx;
        ud[i * nx + j] = pf[(i + top) * pnx_f + j + npml];
        ud[(i + 3) * nx + j] = pf[(nz - i - 1 + top) * pnx_f + j + npml];
    }
    if (t < 3 * nz) {
        const int i = t / 3, j = t - i * 3;
        lr[i * 6 + j] = pf[(i + top) * pnx_f + j + npml];
        lr[i * 6 + j + 3] = pf[(i + top) * pnx_f + nx + npml - j - 1];
    }
}

__global__ void k_ref_read_bdr_ud(const float *ud, float *pb, int nx, int nz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * nx) {
        const int i = t / nx, j = t - i * nx;
        pb[i * nx + j] = ud[i * nx + j];
        pb[(nz - i - 1) * nx + j] = ud[(i + 3) * nx + j];
    }
}

__global__ void k_ref_read_bdr_lr(const float *lr, float *pb, int nx, int nz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * nz) {
        const int i = t / 3, j = t - i * 3;
        pb[i * nx + j] = lr[i * 6 + j];
        pb[i * nx + nx - j - 1] = lr[i * 6 + j + 3];
    }
}

// ============================================================================
// checkpointed backward pass: snapshots and the image from a recomputed segment
// ============================================================================
// max |x - y| (y == nullptr: max |x|) of each of nitem blocks of n floats,
// into bits[item] as the bit pattern of a non-negative float (ordered like
// the value, so an unsigned atomicMax finds the maximum); bits start at 0
__global__ void k_absmax(const float *__restrict__ x, const float *__restrict__ y, long long n, int nitem,
                         unsigned *bits)
{
    const int item = blockIdx.y;
    if (item >= nitem) return;
    const float *xi = x + (long long)item * n;
    const float *yi = y ? y + (long long)item * n : nullptr;
    float m = 0.f;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(yi ? xi[j] - yi[j] : xi[j]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float wm[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) wm[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmaxf(r, wm[w]);
        atomicMax(bits + item, __float_as_uint(r));
    }
}

// BF16 with round-to-nearest-even (finite inputs: the values are scaled into [-1, 1])
__device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// dst[item][j] = (x - y)[item][j] / scale[item] (y == nullptr: x / scale) in
// FP16 (fp16 != 0) or BF16
__global__ void k_pack16(const float *__restrict__ x, const float *__restrict__ y, long long n, int nitem,
                         const unsigned *__restrict__ bits, unsigned short *__restrict__ dst, int fp16)
{
    const int item = blockIdx.y;
    if (item >= nitem) return;
    const float sc = __uint_as_float(bits[item]);
    const float inv = sc > 0.f ? 1.f / sc : 0.f;
    const float *xi = x + (long long)item * n;
    const float *yi = y ? y + (long long)item * n : nullptr;
    unsigned short *di = dst + (long long)item * n;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
        const float v = (yi ? xi[j] - yi[j] : xi[j]) * inv;
        di[j] = fp16 ? __half_as_ushort(__float2half_rn(v)) : f32_to_bf16(v);
    }
}

// x[item][j] = src[item][j] * scale[item], or base - that when base != nullptr
__global__ void k_unpack16(const unsigned short *__restrict__ src, long long n, int nitem,
                           const unsigned *__restrict__ bits, const float *__restrict__ base, float *__restrict__ x,
                           int fp16)
{
    const int item = blockIdx.y;
    if (item >= nitem) return;
    const float sc = __uint_as_float(bits[item]);
    const unsigned short *si = src + (long long)item * n;
    const float *bi = base ? base + (long long)item * n : nullptr;
    float *xi = x + (long long)item * n;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
        const unsigned short h = si[j];
        const float v = (fp16 ? __half2float(__ushort_as_half(h)) : bf16_to_f32(h)) * sc;
        xi[j] = bi ? bi[j] - v : v;
    }
}

// the image of step k from a recomputed segment: for each physical cell and
// shot group, gacc[grp] += sum over the group's live shots (FP64, fixed
// order) of mu^{k+1} X^k, mu^{k+1} = M[(k+1) & 1]
__global__ void k_img_seg(const FwiRecArgs *__restrict__ Ap, const float *__restrict__ xk, int k, int gsize,
                          int ngr)
{
    const FwiRecArgs &A = *Ap;
    const FwiGeomDev &g = A.g;
    const long long nb = (long long)g.nx * g.nz;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int grp = blockIdx.y;
    if (cell >= nb || grp >= ngr) return;
    if (!fwi_image_step(A.img_jump, A.img_k0, A.img_k1, k)) return;
    const int i = (int)(cell / g.nx), j = (int)(cell - (long long)i * g.nx);
    const float *mu = (k & 1) ? A.M0 : A.M1;
    const long long fo = (long long)i * g.ld + j + g.npml;
    double acc = 0.0;
    const int s1 = min((grp + 1) * gsize, A.b.nshots);
    for (int s = grp * gsize; s < s1; ++s) {
        if (k >= A.b.shots[s].nt) continue;
        acc += (double)mu[(long long)s * A.b.fs + fo] * (double)xk[(long long)s * nb + cell];
    }
    gbl(A.gacc)[(long long)grp * nb + cell] += acc;
}

// ============================================================================
// host-side launchers (C linkage inside the library)
// ============================================================================
// The interior and general launches of a step are independent (disjoint
// cells, the same inputs).  With a second stream they run side by side: the
// general strips (latency bound) then overlap the interior ones (memory
// bound).  fork/join events order them against the neighbouring steps.
template <class LI, class LG>
static hipError_t launch_pair(const FwiPair &p, bool has_int, bool has_gen, LI launch_int, LG launch_gen)
{
    if (p.st2 && has_int && has_gen) {
        hipError_t e = hipEventRecord(p.fork, p.st);
        if (e == hipSuccess) e = hipStreamWaitEvent(p.st2, p.fork, 0);
        if (e != hipSuccess) return e;
        launch_gen(p.st2);
        launch_int(p.st);
        e = hipEventRecord(p.join, p.st2);
        if (e == hipSuccess) e = hipStreamWaitEvent(p.st, p.join, 0);
        if (e != hipSuccess) return e;
    } else {
        if (has_int) launch_int(p.st);
        if (has_gen) launch_gen(p.st);
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_fwd_tile_mode(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *const *tl,
                                       const int *ntl, const int *kbase, int koff, const FwiPair &p)
{
    return launch_pair(
        p, ntl[0] > 0, ntl[1] > 0,
        [&](hipStream_t st) {
            hipLaunchKernelGGL((k_fwd_tile<false, MODE>), dim3(((ntl[0] + 31) / 32) * 8), dim3(256), p.lds_int, st, a_dev,
                               a->b.shots, tl[0], ntl[0], kbase, koff);
        },
        [&](hipStream_t st) {
            hipLaunchKernelGGL((k_fwd_tile<true, MODE>), dim3(((ntl[1] + 31) / 32) * 8), dim3(256), 0, st, a_dev,
                               a->b.shots, tl[1], ntl[1], kbase, koff);
        });
}

}  // namespace FWI_KNS
using namespace FWI_KNS;

extern "C" {

#ifndef FWI_SECONDARY
hipError_t fwi_launch_fwd(const FwiFwdArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_fwd_step, dim3(((nblocks + 7) / 8) * 8), dim3(FWI_TX * a->b.sg), 0, st, *a, kbase, koff);
    return hipGetLastError();
}

// tile lists: [0] interior tiles, [1] general tiles (device arrays and sizes)
#endif

hipError_t FWI_VNAME(fwi_launch_fwd_tile)(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *const *tl,
                                          const int *ntl, const int *kbase, int koff, const FwiPair &p, int mode)
{
    if (mode == FWI_FWD_TRACES) return launch_fwd_tile_mode<FWI_FWD_TRACES>(a, a_dev, tl, ntl, kbase, koff, p);
    if (mode == FWI_FWD_XSEG) return launch_fwd_tile_mode<FWI_FWD_XSEG>(a, a_dev, tl, ntl, kbase, koff, p);
    if (mode == FWI_FWD_PLAIN) return launch_fwd_tile_mode<FWI_FWD_PLAIN>(a, a_dev, tl, ntl, kbase, koff, p);
    return launch_fwd_tile_mode<FWI_FWD_FRAMES>(a, a_dev, tl, ntl, kbase, koff, p);
}

// the forward step as one launch (k_fwd_mixed) over the work order `work` of
// nwork workgroups (a multiple of 8)
hipError_t FWI_VNAME(fwi_launch_fwd_mixed)(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *work, int nwork,
                                           const int *const *tl, const int *ntl, const int *kbase, int koff,
                                           hipStream_t st, int mode)
{
    if (nwork <= 0) return hipSuccess;
#define FWI_MIXED(M)                                                                                            \
    hipLaunchKernelGGL(k_fwd_mixed<M>, dim3(nwork), dim3(256), 0, st, a_dev, a->b.shots, work, tl[0], ntl[0], \
                       tl[1], ntl[1], kbase, koff)
    if (mode == FWI_FWD_TRACES) FWI_MIXED(FWI_FWD_TRACES);
    else if (mode == FWI_FWD_XSEG) FWI_MIXED(FWI_FWD_XSEG);
    else if (mode == FWI_FWD_PLAIN) FWI_MIXED(FWI_FWD_PLAIN);
    else FWI_MIXED(FWI_FWD_FRAMES);
#undef FWI_MIXED
    return hipGetLastError();
}

// the fused backward step as one launch (k_bwd_mixed) over the work order
// `work` of nwork workgroups (a multiple of 8); tl[0..3] as for
// fwi_launch_bwd_tile, the reconstruction items imaging one step late
hipError_t FWI_VNAME(fwi_launch_bwd_mixed)(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                           const int *work, int nwork, const int *const *tl, const int *ntl,
                                           const int *kbase, int koff, int gsize, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bwd_mixed, dim3(nwork), dim3(256), 0, st, a_dev, r_dev, a->b.shots, work, tl[0], tl[1],
                       ntl[1], tl[2], ntl[2], tl[3], kbase, koff, gsize);
    return hipGetLastError();
}

hipError_t FWI_VNAME(fwi_launch_adj_tile)(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const int *const *tl,
                                          const int *ntl, const int *kbase, int koff, const FwiPair &p)
{
    return launch_pair(
        p, ntl[0] > 0, ntl[1] > 0,
        [&](hipStream_t st) {
            hipLaunchKernelGGL(k_adj_tile<false>, dim3(((ntl[0] + 31) / 32) * 8), dim3(256), 0, st, a_dev,
                               a->b.shots, tl[0], ntl[0], kbase, koff);
        },
        [&](hipStream_t st) {
            hipLaunchKernelGGL(k_adj_tile<true>, dim3(((ntl[1] + 31) / 32) * 8), dim3(256), 0, st, a_dev,
                               a->b.shots, tl[1], ntl[1], kbase, koff);
        });
}

hipError_t FWI_VNAME(fwi_launch_rec_tile)(const FwiRecArgs *a, const FwiRecArgs *a_dev, const int *const *tl,
                                          const int *ntl, const int *kbase, int koff, int gsize, const FwiPair &p)
{
    return launch_pair(
        p, ntl[0] > 0, ntl[1] > 0,
        [&](hipStream_t st) {
            hipLaunchKernelGGL(k_rec_tile_int, dim3(((ntl[0] + 7) / 8) * 8), dim3(256), 0, st, a_dev, a->b.shots,
                               tl[0], ntl[0], kbase, koff, gsize);
        },
        [&](hipStream_t st) {
            hipLaunchKernelGGL(k_rec_tile_gen, dim3(((ntl[1] + 7) / 8) * 8), dim3(256), 0, st, a_dev, a->b.shots,
                               tl[1], ntl[1], kbase, koff, gsize);
        });
}

// The two sides of a fused backward step.  The interior side: the fused items
// (tl[0]) with the occupancy reserve lds.  The general side: every adjoint
// strip of the other tiles (tl[1]: the light strips, tl[2]: the corner
// strips, one launch of k_adj_tile_gen), then the reconstruction of those
// tiles (tl[3]), which reads their mu^{k+1}, on the same stream.  The general
// side may come from the build with taller strips than the interior side's:
// the host then builds its lists in that build's strip height.
hipError_t FWI_VNAME(fwi_launch_bwd_int)(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                         const int *tl0, int n0, const int *kbase, int koff, int gsize, int lds,
                                         hipStream_t st)
{
    if (n0 > 0)
        hipLaunchKernelGGL(k_bwd_tile_int, dim3(((n0 + 7) / 8) * 8), dim3(256), lds, st, a_dev, r_dev, a->b.shots,
                           tl0, n0, kbase, koff, gsize);
    return hipGetLastError();
}

hipError_t FWI_VNAME(fwi_launch_bwd_gen)(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                         const int *const *tl, const int *ntl, const int *kbase, int koff, int gsize,
                                         hipStream_t st)
{
    if (ntl[1] > 0 || ntl[2] > 0)
        hipLaunchKernelGGL(k_adj_tile_gen, dim3(fwi_gen_blocks(ntl[2]) + fwi_gen_blocks(ntl[1])), dim3(256), 0, st,
                           a_dev, a->b.shots, tl[2], ntl[2], tl[1], ntl[1], kbase, koff);
    if (ntl[3] > 0)
        hipLaunchKernelGGL(k_rec_tile_pad, dim3(((ntl[3] + 7) / 8) * 8), dim3(256), 0, st, r_dev, a->b.shots, tl[3],
                           ntl[3], kbase, koff, gsize);
    return hipGetLastError();
}

// one backward step of this build's strip height: the interior side on p.st;
// with p.st2 the general side runs on st2 beside it (the two sides write
// disjoint cells of mu^{k+1}, p^{k-1} and the image and read only the other
// side's previous-step fields).  The general side's two launches share a
// stream, so nothing inside the step waits for another stream (a wait across
// streams costs about 10 us here).
hipError_t FWI_VNAME(fwi_launch_bwd_tile)(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                          const int *const *tl, const int *ntl, const int *kbase, int koff,
                                          int gsize, const FwiPair &p)
{
    const bool has_gen = ntl[1] > 0 || ntl[2] > 0 || ntl[3] > 0;
    return launch_pair(
        p, ntl[0] > 0, has_gen,
        [&](hipStream_t st) {
            (void)FWI_VNAME(fwi_launch_bwd_int)(a, a_dev, r_dev, tl[0], ntl[0], kbase, koff, gsize, p.lds_int, st);
        },
        [&](hipStream_t st) {
            (void)FWI_VNAME(fwi_launch_bwd_gen)(a, a_dev, r_dev, tl, ntl, kbase, koff, gsize, st);
        });  // launch errors: hipGetLastError at the end of launch_pair
}

#ifndef FWI_SECONDARY
hipError_t fwi_launch_adj(const FwiAdjArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_step, dim3(((nblocks + 7) / 8) * 8), dim3(FWI_TX * a->b.sg), 0, st, *a, kbase, koff);
    return hipGetLastError();
}

hipError_t fwi_launch_rec(const FwiRecArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_rec_step, dim3(((nblocks + 7) / 8) * 8), dim3(FWI_TX * a->b.sg), 0, st, *a, kbase, koff);
    return hipGetLastError();
}

hipError_t fwi_launch_advance(int *kbase, int delta, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, kbase, delta);
    return hipGetLastError();
}

hipError_t fwi_launch_set_int(int *p, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, v);
    return hipGetLastError();
}

hipError_t fwi_launch_fill_f32(float *p, long long n, float v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t fwi_launch_fill_f64(double *p, long long n, double v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t fwi_launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, st, in, out, rows, cols);
    return hipGetLastError();
}

hipError_t fwi_launch_misfit_l2(const float *syn, const float *obs, float *res, long long n, double *partial,
                                int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_misfit_l2, dim3(nblocks), dim3(256), 0, st, syn, obs, res, n, partial);
    return hipGetLastError();
}

hipError_t fwi_launch_grad_finalize(const double *gacc, int ngroups, long long n, const float *vel, double *gk,
                                    double *grad, hipStream_t st)
{
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_grad_finalize, dim3((unsigned)nb), dim3(256), 0, st, gacc, ngroups, n, vel, gk, grad);
    return hipGetLastError();
}

hipError_t fwi_launch_ref_phase(const FwiRefArgs *a, int phase, hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grid((a->pnx + 2 + 63) / 64, (a->pnz + 2 + 3) / 4);
    hipLaunchKernelGGL(k_ref_phase, grid, blk, 0, st, *a, phase);
    return hipGetLastError();
}

hipError_t fwi_launch_ref_time_integral(float *pnew, const float *pnow, const float *pold, const float *pxx,
                                        const float *pzz, const float *ekp, long long n, hipStream_t st)
{
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_ref_time_integral, dim3((unsigned)nb), dim3(256), 0, st, pnew, pnow, pold, pxx, pzz, ekp, n);
    return hipGetLastError();
}

hipError_t fwi_launch_ref_write_bdr(float *ud, float *lr, const float *pf, int nx, int nz, int npml, int top,
                                    hipStream_t st)
{
    const int n = 3 * (nx > nz ? nx : nz);
    hipLaunchKernelGGL(k_ref_write_bdr, dim3((n + 255) / 256), dim3(256), 0, st, ud, lr, pf, nx, nz, npml, top);
    return hipGetLastError();
}

hipError_t fwi_launch_ref_read_bdr(const float *ud, const float *lr, float *pb, int nx, int nz, hipStream_t st)
{
    hipLaunchKernelGGL(k_ref_read_bdr_ud, dim3((3 * nx + 255) / 256), dim3(256), 0, st, ud, pb, nx, nz);
    hipLaunchKernelGGL(k_ref_read_bdr_lr, dim3((3 * nz + 255) / 256), dim3(256), 0, st, lr, pb, nx, nz);
    return hipGetLastError();
}


hipError_t fwi_launch_absmax(const float *x, const float *y, long long n, int nitem, unsigned *bits, hipStream_t st)
{
    if (nitem <= 0 || n <= 0) return hipSuccess;
    // few blocks per item, each reducing a long stretch first: the items'
    // maxima sit side by side in one cache line, so every block's atomicMax
    // on them is serialised (1024 blocks per item made a 16-bit snapshot of
    // 64 shots take about 0.5 ms)
    const long long per = (n + 255) / 256;
    const int bx = (int)std::min<long long>(per, 32);
    hipLaunchKernelGGL(k_absmax, dim3(bx, nitem), dim3(256), 0, st, x, y, n, nitem, bits);
    return hipGetLastError();
}

hipError_t fwi_launch_pack16(const float *x, const float *y, long long n, int nitem, const unsigned *bits,
                             unsigned short *dst, int fp16, hipStream_t st)
{
    if (nitem <= 0 || n <= 0) return hipSuccess;
    const int bx = (int)std::min<long long>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_pack16, dim3(bx, nitem), dim3(256), 0, st, x, y, n, nitem, bits, dst, fp16);
    return hipGetLastError();
}

hipError_t fwi_launch_unpack16(const unsigned short *src, long long n, int nitem, const unsigned *bits,
                               const float *base, float *x, int fp16, hipStream_t st)
{
    if (nitem <= 0 || n <= 0) return hipSuccess;
    const int bx = (int)std::min<long long>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_unpack16, dim3(bx, nitem), dim3(256), 0, st, src, n, nitem, bits, base, x, fp16);
    return hipGetLastError();
}

hipError_t fwi_launch_img_seg(const FwiRecArgs *a, const FwiRecArgs *a_dev, const float *xk, int k, int gsize,
                              int ngr, hipStream_t st)
{
    const long long nb = (long long)a->g.nx * a->g.nz;
    hipLaunchKernelGGL(k_img_seg, dim3((unsigned)((nb + 255) / 256), ngr), dim3(256), 0, st, a_dev, xk, k, gsize,
                       ngr);
    return hipGetLastError();
}

#endif  // FWI_SECONDARY

}  // extern "C"
