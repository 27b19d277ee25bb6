// fwi_dev.h -- launch parameter blocks shared by the HIP kernels
// (fwi_kernels.hip) and the C-ABI host layer (fwi_hip.cpp).
//
// Device layout of one batch (all FP32 unless noted):
//   f-grid fields  [shot][pnz][ld]    ld = 64 * ceil(pnx / 64), padding zero
//   p^k of a shot lives in buffer P[k & 1]; the update writes p^{k+1} over
//   p^{k-1} (two buffers, in place).  The adjoint mu^m lives in M[m & 1].
//   CPML memory (forward psi, adjoint chi), per shot:
//     xh  [2][pnz][2*(npml+1)]  half (u) nodes, left strip then right strip
//     xd  [2][pnz][2*npml]      p nodes of the x strips
//     zh  [2][npml+1][ld]       half (v) nodes of the bottom strip
//     zd  [2][npml][ld]         p nodes of the bottom strip
//   ping-pong buffer index = step & 1 (read old, write new).
//   frame strips [shot][nt][6nx+6nz]: rows {0,1,2,nz-1,nz-2,nz-3} x nx, then
//   nz x cols {0,1,2,nx-1,nx-2,nx-3} (the bc_ud / bc_lr order of
//   write_bdr, fwi_2d.c:893-910), one block per step.
//   traces: time-major [shot][nt][nr] while propagating.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FWI_TX 64      // x cells per wave (one lane per column)
// lane byte offset of a buffer load outside every field: the buffer
// resources of the tile kernels cover one shot's field (or less), which the
// host keeps below this size, so the resource's range check returns 0
#define FWI_BUF_OUT 0x40000000
#define FWI_MAXSG 4    // shots (waves) per workgroup

struct FwiShotDev {
    int nt;          // time steps of this shot
    int isx, isz;    // source cell on the f-grid (x already offset by npml)
    int nr;          // receivers
    long long src_off;    // into srcf: wav[it] * iarea, nt entries
    long long trace_off;  // into syn/res (time-major nt x nr)
    long long strip_off;  // into pstrip/xstrip (nt x sl)
    int rrow_off;         // into rrow (pnz entries: receiver-row slot or -1)
    int rz_lo, rz_hi;     // rows holding receivers: [rz_lo, rz_hi] (empty if lo > hi)
    int pad_;
};

struct FwiGeomDev {
    int nx, nz, npml, pnx, pnz, ld;
    int tz;              // tile height in rows (>= 8)
    int ntx, ntz;        // f-grid tiles
    int bntx, bntz;      // physical-grid tiles
    int sl;              // strip length per step = 6 nx + 6 nz
    float c0x, c1x, ix, c0z, c1z, iz, iz840, iz24;
    int ntz2;            // f-grid tiles of the 2-D tile kernels (FWI_TZ2 rows each)
    int pad_;
};

// Rows of a 2-D tile (4 waves x FWI_TZ2/4 output rows).  The tile kernels are
// compiled twice: with 8-row strips (the default) and with 4-row strips (the
// object built with FWI_TZ2 = 16, its launchers suffixed _s4); a context picks
// one per pass (fwi_hip_set_pass_shape).  Host code passes the strip height
// to the functions below; device code uses its own FWI_STRIP.
#ifndef FWI_TZ2
#define FWI_TZ2 32
#endif

// rows per wave of the tile reconstruction kernel (its FP64 image
// accumulators grow with it, so it is set apart from the forward/adjoint);
// the fused backward step needs it equal to the strip height
#ifndef FWI_REC_ROWS
#define FWI_REC_ROWS 8
#endif

// Strip classes of the 2-D tile kernels.  A wave owns a strip of 64 columns x
// FWI_STRIP rows; a strip is "interior" when none of its cells, nor any u/v
// node it evaluates, needs a CPML, edge, free-surface, frame, source or
// receiver rule.  Interior and general strips run in separate launches (each
// with its own register budget); the host builds the lists with these same
// functions.
#define FWI_STRIP (FWI_TZ2 / 4)
// forward strip at (x0, r0)
__host__ __device__ inline bool fwi_fwd_strip_interior(const FwiGeomDev &g, const FwiShotDev &sh, int x0, int r0,
                                                      int fr = FWI_STRIP)
{
    const int r1 = r0 + fr;
    if (x0 < g.npml + 3 || x0 + 66 > g.pnx - g.npml - 1) return false;  // u nodes x0-1 .. x0+65
    if (r0 < 3 || r1 > g.nz - 3) return false;                           // v nodes r0-1 .. r1+1
    if (sh.isz >= r0 && sh.isz < r1 && sh.isx >= x0 && sh.isx < x0 + FWI_TX) return false;
    if (sh.rz_lo <= sh.rz_hi && sh.rz_lo < r1 && sh.rz_hi >= r0) return false;
    return true;
}
// Special rules a general strip can meet.  Each general body is compiled for
// a mask of them (the rest of its code is left out); a strip runs in the
// smallest compiled mask holding its class.  The conditions are those the
// bodies test at run time (fwd_tile_rows / adj_tile_rows).
#define FWI_CLS_X 1    // x CPML strips or grid edges within the x stencil
#define FWI_CLS_Z 2    // free-surface rows, bottom CPML rows or the last rows
#define FWI_CLS_RS 4   // receiver rows or the source cell
#define FWI_CLS_FR 8   // frame rows or columns (forward)
#define FWI_CLS_ALL 15
__host__ __device__ inline int fwi_fwd_strip_class(const FwiGeomDev &g, const FwiShotDev &sh, int x0, int r0,
                                                   int fr = FWI_STRIP)
{
    const int npml = g.npml, FR = fr;
    int c = 0;
    if (!((x0 >= npml + 3) && (x0 + 66 <= g.pnx - npml - 1))) c |= FWI_CLS_X;
    if (!(r0 >= 3 && r0 + FR + 1 < g.pnz - npml)) c |= FWI_CLS_Z;
    if ((sh.rz_lo <= sh.rz_hi && sh.rz_lo < r0 + FR && sh.rz_hi >= r0) ||
        (sh.isz >= r0 && sh.isz < r0 + FR && sh.isx >= x0 && sh.isx < x0 + FWI_TX))
        c |= FWI_CLS_RS;
    if ((r0 < 3 || (r0 + FR > g.nz - 3 && r0 < g.nz)) ||
        ((x0 < npml + 3 && x0 + FWI_TX > npml) || (x0 < npml + g.nx && x0 + FWI_TX > npml + g.nx - 3)))
        c |= FWI_CLS_FR;
    return c;
}
__host__ __device__ inline int fwi_adj_strip_class(const FwiGeomDev &g, const FwiShotDev &sh, int x0, int r0,
                                                   int fr = FWI_STRIP)
{
    const int npml = g.npml, FR = fr;
    int c = 0;
    if (!((x0 >= npml + 3) && (x0 + 66 <= g.pnx - npml - 1))) c |= FWI_CLS_X;
    if (!(r0 >= 6 && r0 + FR + 2 < g.pnz - npml)) c |= FWI_CLS_Z;
    if (sh.rz_lo <= sh.rz_hi && sh.rz_lo < r0 + FR && sh.rz_hi >= r0) c |= FWI_CLS_RS;
    return c;
}

// adjoint strip: no source term; the transposed free-surface rows reach node 4
__host__ __device__ inline bool fwi_adj_strip_interior(const FwiGeomDev &g, const FwiShotDev &sh, int x0, int r0,
                                                      int fr = FWI_STRIP)
{
    const int r1 = r0 + fr;
    if (x0 < g.npml + 3 || x0 + 66 > g.pnx - g.npml - 1) return false;
    if (r0 < 6 || r1 > g.nz - 3) return false;
    if (sh.rz_lo <= sh.rz_hi && sh.rz_lo < r1 && sh.rz_hi >= r0) return false;
    return true;
}
// reconstruction tile: 64 x FWI_REC_ROWS physical cells at (bx0, z0), every
// cell inner (3 or more from the edge) and the source elsewhere
__host__ __device__ inline bool fwi_rec_tile_interior(const FwiGeomDev &g, const FwiShotDev &sh, int bx0, int z0,
                                                     int rr = FWI_REC_ROWS)
{
    if (bx0 < 3 || bx0 + 66 >= g.nx) return false;
    if (z0 < 3 || z0 + rr > g.nz - 3) return false;
    const int sx = sh.isx - g.npml;
    if (sh.isz >= z0 && sh.isz < z0 + rr && sx >= bx0 && sx < bx0 + FWI_TX) return false;
    return true;
}

struct FwiBatchDev {
    int nshots, sg, ngroups;
    int pad_;
    const FwiShotDev *shots;
    long long fs;        // field elements per shot (pnz * ld)
    long long xh_s, xd_s, zh_s, zd_s;   // per-shot stride of the CPML arrays
    long long xh_b, xd_b, zh_b, zd_b;   // ping-pong buffer stride
    const float *ekp;    // [pnz][ld]
    const float *tab;    // 6 x (2 npml + 1): a0x b0x a1x b1x a1z b1z
    const float *srcf;
    const int *rrow;
    const int *rmap;     // [slot][ld]
};

struct FwiFwdArgs {
    FwiGeomDev g;
    FwiBatchDev b;
    float *P0, *P1;
    float *xh, *xd, *zh, *zd;
    float *syn;
    float *pstrip, *xstrip;
    // segment recompute of the checkpointed backward pass: X^k of every
    // physical cell at xseg[(k - xseg_k0) * xseg_step + s nx nz + i nx + j]
    float *xseg;
    long long xseg_step;
    int xseg_k0, pad2_;
};

// forward tile kernel modes
#define FWI_FWD_FRAMES 0   // record traces, save the boundary frames (boundary-saving reconstruction)
#define FWI_FWD_TRACES 1   // record traces only (checkpointed forward pass)
#define FWI_FWD_XSEG 2     // store X on the physical grid only (segment recompute)
#define FWI_FWD_PLAIN 3    // record nothing (advancing a restored state)

struct FwiAdjArgs {
    FwiGeomDev g;
    FwiBatchDev b;
    float *M0, *M1;
    float *xh, *xd, *zh, *zd;
    const float *res;
};

struct FwiRecArgs {
    FwiGeomDev g;
    FwiBatchDev b;
    float *P0, *P1;
    const float *M0, *M1;
    const float *pstrip, *xstrip;
    double *gacc;        // [ngroups][nz][nx]
    int img_jump, img_k0, img_k1;  // steps imaged: k0 <= k <= k1, (k - k0) % jump == 0
    int pad_;
    // X^k of the general tiles of the one-launch backward step, which image
    // one step late: xg[(k & 1) * xg_b + s nx nz + i nx + j]
    float *xg;
    long long xg_b;
};

// whether reconstruction step k adds to the image
__host__ __device__ inline bool fwi_image_step(int jump, int k0, int k1, int k)
{
    return k >= k0 && k <= k1 && (k - k0) % jump == 0;
}

// streams of a step's interior and general launches (st2 = nullptr: both on st)
struct FwiPair {
    hipStream_t st, st2;
    hipEvent_t fork, join;
    int lds_int;  // dynamic LDS bytes reserved per interior workgroup (occupancy cap)
};
// workgroups of k_adj_tile_gen for n strips (4 per workgroup), a multiple of 8
__host__ __device__ inline int fwi_gen_blocks(int n) { return ((((n + 3) >> 2) + 7) >> 3) << 3; }
// LDS of one workgroup of the tile kernels with fr-row strips: interior (wave
// rows) and general (wave rows + CPML tables)
__host__ __device__ constexpr int fwi_lds_tile(int fr) { return 4 * 2 * fr * 72 * 4; }
__host__ __device__ constexpr int fwi_lds_gen(int fr) { return fwi_lds_tile(fr) + 6 * (2 * 64 + 1) * 4; }

// per-step kernels with the reference's loop structure (unit parity)
struct FwiRefArgs {
    int pnx, pnz, npml;
    float c0x, c1x, ix, c0z, c1z, iz, iz840, iz24;
    const float *p;
    float *u, *v, *ux, *vz;
    float *u0, *u1, *v1, *ux0, *ux1, *vz1;
    const float *tab;
};
