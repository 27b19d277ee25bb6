// fwi_hip.cpp -- host side of the C ABI declared in include/fwi_hip.h.
//
// Owns the device buffers of a batch of shots, prepares the launch
// parameter blocks of fwi_kernels.hip and sequences the time loops.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <climits>
#include <vector>

#include "fwi_dev.h"
#include "../../include/fwi_hip.h"

extern "C" {
hipError_t fwi_launch_fwd(const FwiFwdArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st);
hipError_t fwi_launch_fwd_tile(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *const *tl, const int *ntl,
                               const int *kbase, int koff, const FwiPair &p, int mode);
hipError_t fwi_launch_absmax(const float *x, const float *y, long long n, int nitem, unsigned *bits, hipStream_t st);
hipError_t fwi_launch_pack16(const float *x, const float *y, long long n, int nitem, const unsigned *bits,
                             unsigned short *dst, int fp16, hipStream_t st);
hipError_t fwi_launch_unpack16(const unsigned short *src, long long n, int nitem, const unsigned *bits,
                               const float *base, float *x, int fp16, hipStream_t st);
hipError_t fwi_launch_img_seg(const FwiRecArgs *a, const FwiRecArgs *a_dev, const float *xk, int k, int gsize,
                              int ngr, hipStream_t st);
hipError_t fwi_launch_adj(const FwiAdjArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st);
hipError_t fwi_launch_adj_tile(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const int *const *tl, const int *ntl,
                               const int *kbase, int koff, const FwiPair &p);
hipError_t fwi_launch_rec_tile(const FwiRecArgs *a, const FwiRecArgs *a_dev, const int *const *tl, const int *ntl,
                               const int *kbase, int koff, int gsize, const FwiPair &p);
hipError_t fwi_launch_rec(const FwiRecArgs *a, const int *kbase, int koff, int nblocks, hipStream_t st);
hipError_t fwi_launch_bwd_tile(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                               const int *const *tl, const int *ntl, const int *kbase, int koff, int gsize,
                               const FwiPair &p);
hipError_t fwi_launch_fwd_mixed(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *work, int nwork,
                                const int *const *tl, const int *ntl, const int *kbase, int koff, hipStream_t st,
                                int mode);
hipError_t fwi_launch_fwd_mixed_s4(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *work, int nwork,
                                   const int *const *tl, const int *ntl, const int *kbase, int koff, hipStream_t st,
                                   int mode);
hipError_t fwi_launch_bwd_mixed(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                const int *work, int nwork, const int *const *tl, const int *ntl, const int *kbase,
                                int koff, int gsize, hipStream_t st);
hipError_t fwi_launch_bwd_mixed_s4(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                   const int *work, int nwork, const int *const *tl, const int *ntl, const int *kbase,
                                   int koff, int gsize, hipStream_t st);
// the same tile kernels built with 4-row strips
hipError_t fwi_launch_fwd_tile_s4(const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *const *tl,
                                  const int *ntl, const int *kbase, int koff, const FwiPair &p, int mode);
hipError_t fwi_launch_adj_tile_s4(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const int *const *tl,
                                  const int *ntl, const int *kbase, int koff, const FwiPair &p);
hipError_t fwi_launch_rec_tile_s4(const FwiRecArgs *a, const FwiRecArgs *a_dev, const int *const *tl,
                                  const int *ntl, const int *kbase, int koff, int gsize, const FwiPair &p);
hipError_t fwi_launch_bwd_tile_s4(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                  const int *const *tl, const int *ntl, const int *kbase, int koff, int gsize,
                                  const FwiPair &p);
// the two sides of a fused backward step (fwi_launch_bwd_tile), apart: the
// interior side of the 4-row build beside the general side of the 8-row one
hipError_t fwi_launch_bwd_int_s4(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                                 const int *tl0, int n0, const int *kbase, int koff, int gsize, int lds,
                                 hipStream_t st);
hipError_t fwi_launch_bwd_gen(const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                              const int *const *tl, const int *ntl, const int *kbase, int koff, int gsize,
                              hipStream_t st);
hipError_t fwi_launch_advance(int *kbase, int delta, hipStream_t st);
hipError_t fwi_launch_set_int(int *p, int v, hipStream_t st);
hipError_t fwi_launch_fill_f32(float *p, long long n, float v, hipStream_t st);
hipError_t fwi_launch_fill_f64(double *p, long long n, double v, hipStream_t st);
hipError_t fwi_launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st);
hipError_t fwi_launch_misfit_l2(const float *syn, const float *obs, float *res, long long n, double *partial,
                                int nblocks, hipStream_t st);
hipError_t fwi_launch_grad_finalize(const double *gacc, int ngroups, long long n, const float *vel, double *gk,
                                    double *grad, hipStream_t st);
}

extern "C" {
hipError_t fwi_launch_ref_phase(const FwiRefArgs *a, int phase, hipStream_t st);
hipError_t fwi_launch_ref_time_integral(float *pnew, const float *pnow, const float *pold, const float *pxx,
                                        const float *pzz, const float *ekp, long long n, hipStream_t st);
hipError_t fwi_launch_ref_write_bdr(float *ud, float *lr, const float *pf, int nx, int nz, int npml, int top,
                                    hipStream_t st);
hipError_t fwi_launch_ref_read_bdr(const float *ud, const float *lr, float *pb, int nx, int nz, hipStream_t st);
}

static char g_err[512] = "";

// ---------------------------------------------------------------------------
// Binomial checkpointing of the backward pass (the "revolve" schedule of
// Griewank and Walther, ACM TOMS 26 (2000) 19-45).
//
// The backward pass needs the forward states S_k (p^k, p^{k-1} and the CPML
// memory before forward step k) for k = t-1 down to 0.  A turn at k is one
// forward step from S_k that gives X^k, then the adjoint step k and its
// imaging.  With s snapshot slots, one of them holding S_a (the live state),
// steps [a, b) are reversed as: advance j steps to S_{a+j}, snapshot it,
// reverse [a+j, b) with s-1 slots, restore S_a, reverse [a, a+j) with s
// slots; with one slot, restore S_a and advance to each step in turn.
// beta(s, r) = C(s + r, s) steps can be reversed with s slots and no step
// swept forward more than r times.  For l = b - a in (beta(s, r-1),
// beta(s, r)] the splits j in [max(beta(s, r-2), l - beta(s-1, r)),
// min(beta(s, r-1), l - beta(s-1, r-1))] (beta(s, -1) = 0) reach the least
// number of advance steps, r l - beta(s+1, r-1); the schedule takes the
// smallest.
// ---------------------------------------------------------------------------
enum { RV_ADVANCE = 0, RV_TAKESHOT = 1, RV_RESTORE = 2, RV_TURN = 3 };
struct RvAct {
    int op, a, b, slot;  // advance S_a -> S_b; snapshot / restore S_a in slot; turn at a
};

static long long rv_beta(long long s, long long r)
{
    const long long cap = 1LL << 61;
    long long v = 1;  // C(s + i, i) for i = 0 .. r
    for (long long i = 1; i <= r; i++) {
        if (v > cap / (s + i)) return cap;
        v = v * (s + i) / i;
    }
    return v;
}

static void rv_reverse(int a, int b, int slot_a, int s, int next_slot, std::vector<RvAct> &out)
{
    while (b > a) {
        if (b - a == 1) {
            out.push_back({RV_TURN, a, a + 1, -1});
            return;
        }
        if (s <= 1) {  // S_a only: advance to every step from it, the last step first
            for (int k = b - 1; k >= a; k--) {
                if (k < b - 1) out.push_back({RV_RESTORE, a, a, slot_a});
                if (k > a) out.push_back({RV_ADVANCE, a, k, -1});
                out.push_back({RV_TURN, k, k + 1, -1});
            }
            return;
        }
        const long long l = b - a;
        long long r = 0;
        while (rv_beta(s, r) < l) r++;
        const long long lo = std::max(r >= 2 ? rv_beta(s, r - 2) : 0LL, l - rv_beta(s - 1, r));
        const int j = (int)std::max(1LL, lo);
        out.push_back({RV_ADVANCE, a, a + j, -1});
        out.push_back({RV_TAKESHOT, a + j, a + j, next_slot});
        rv_reverse(a + j, b, next_slot, s - 1, next_slot + 1, out);
        out.push_back({RV_RESTORE, a, a, slot_a});
        b = a + j;  // then [a, a + j) with the same s slots (tail iteration)
    }
}

// the schedule for t steps with `snaps` slots, S_0 in slot 0
static void rv_schedule(int t, int snaps, std::vector<RvAct> &out)
{
    out.clear();
    if (t >= 1) rv_reverse(0, t, 0, std::max(1, snaps), 1, out);
}

struct fwi_hip_ctx {
    fwi_hip_desc d;
    FwiGeomDev g;
    hipStream_t st = nullptr;
    bool own_stream = false;
    // second stream for the general strips of a step, which can run beside the
    // interior ones (default priority, like st); fork/join events order them.
    // On by default: the general strips are latency bound (long dependent
    // load chains at 4 waves per SIMD) and fill the interior launches' gaps
    // and tails (596 -> 621 shot-gradients/s at 1024x320 x 64 shots, nt=400)
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool concurrent = true;
    // rows per wave of the tile kernels in the forward pass (and the
    // recompute of the checkpointed backward pass) and in the backward pass:
    // 8 or 4 (fwi_hip_set_pass_shape); the strip lists are built for them
    int strip_fwd = 4, strip_bwd = 4;
    // the forward step as one launch of interior and general strips
    // (k_fwd_mixed, fwi_hip_set_fwd_mixed) instead of two streams joined per
    // step; fwd_work holds its workgroup order (n_fwd_work, a multiple of 8)
    bool fwd_mixed = true;
    int *fwd_work = nullptr;
    int n_fwd_work = 0;
    // the fused backward step as one launch (k_bwd_mixed, fwi_hip_set_bwd_mixed);
    // bwd_work holds the workgroup order of every shot chunk, chunk ch at
    // bwd_work_off[ch] with bwd_work_n[ch] entries, built for bwd_work_for
    bool bwd_mixed = false;
    int *bwd_work = nullptr;
    long long bwd_work_cap = 0;
    std::vector<long long> bwd_work_off;
    std::vector<int> bwd_work_n;
    long long bwd_work_for = -1;
    // With the second stream, the interior launches of the forward and fused
    // backward steps reserve unused LDS so that at most int_wgs_* of their
    // workgroups share a CU, leaving room for a general workgroup beside them
    // (0: no cap).  Defaults: 4-row strips with 6 forward and 4 backward
    // workgroups per CU (1024x320 x 64 shots: forward 64.1 -> 61.6 us, backward
    // 151.5 -> 148.0 us against 8-row strips at 3; 4096x1024 x 16 shots:
    // forward 199.8 -> 184.7 us, backward unchanged)
    int int_wgs_fwd = 6, int_wgs_bwd = 4;
    int lds_cu = 0;  // LDS bytes per CU of the device
    // backward pass: adjoint step k on st beside imaging step k+1 on st2
    // (independent: the adjoint writes only the mu buffer the imaging step
    // does not read, and never touches p or the images)
    bool bwd_overlap = false;
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    // the second chunk in flight of the backward pass (bwd_lanes 2): its
    // interior and general streams, their fork/join events, and the events
    // that hand it over from st at the start of a graph block and back at
    // the end
    hipStream_t stB = nullptr, stB2 = nullptr;
    hipEvent_t evB_fork = nullptr, evB_join = nullptr, evL_fork = nullptr, evL_join = nullptr;
    // the checkpointed backward pass images step k on stB beside the adjoint
    // step k - 1: adjoint step k done, imaging steps k (ring of two) done
    hipEvent_t ev_adjd = nullptr, ev_img[2] = {nullptr, nullptr};
    // host <-> device trace copies beside the steps (forward_tm, a residual
    // from set_residual_tm): a copy stream and one event pair per graph block
    hipStream_t stc = nullptr;
    std::vector<hipEvent_t> blk_done, blk_copy;
    const float *res_host = nullptr;  // time-major residual the next adjoint uploads
    char err[512] = "";
    size_t bytes = 0;
    std::vector<std::pair<void *, size_t>> allocs;  // live allocations and their sizes

    // model
    float *ekp = nullptr, *vel = nullptr;
    bool have_model = false;
    // wavefields and CPML memory
    float *P0 = nullptr, *P1 = nullptr, *M0 = nullptr, *M1 = nullptr;
    float *xh = nullptr, *xd = nullptr, *zh = nullptr, *zd = nullptr;
    long long fs = 0, xh_b = 0, xd_b = 0, zh_b = 0, zd_b = 0;
    float *tab = nullptr;
    // frames and traces
    float *pstrip = nullptr, *xstrip = nullptr;  // boundary frames, allocated by the first boundary-saving forward
    // X^k of the general tiles of the one-launch backward step, kept one step
    // for their late imaging (2 x max_shots x nx nz), allocated by the first
    // such backward pass
    float *xg = nullptr;
    // checkpointed backward pass (ckpt_every > 0): snapshots of p every
    // ckpt_every steps in FP32/FP16/BF16 (per-shot max-abs scale for 16 bit)
    // with the CPML memory in FP32, a second CPML set for the recompute,
    // and X of one segment on the physical grid
    int ckpt_every = 0, ckpt_storage = 0;
    void *ck_store = nullptr;
    unsigned *ck_scale = nullptr;
    float *ck_cpml = nullptr, *ck_xseg = nullptr;
    float *xh2 = nullptr, *xd2 = nullptr, *zh2 = nullptr, *zd2 = nullptr;
    long long ck_nseg = 0, ck_K = 0, ck_S = 0;  // capacity of the snapshot buffers
    // binomial checkpointing (set_revolve): rv_snaps slots (-1: from the HBM
    // budget); the backward method the last forward pass prepared (0 frames,
    // 1 uniform snapshots, 2 binomial) with its slots and schedule (the
    // actions before the first turn ran during the forward pass); the
    // forward steps the last adjoint pass recomputed
    int rv_snaps = 0, bwd_method = 0, bwd_snaps = 0;
    std::vector<RvAct> rv_sched;
    int rv_sched_t = -1, rv_sched_s = -1;
    size_t rv_prefix = 0;
    long long recomputed = 0;
    FwiFwdArgs *fwd_rc_d = nullptr;
    FwiFwdArgs fwd_rc_h;
    float *syn_tm = nullptr, *syn = nullptr, *obs = nullptr, *res = nullptr, *res_tm = nullptr;
    double *gacc = nullptr, *gk = nullptr, *grad = nullptr, *partial = nullptr;
    int npartial = 1024;
    // batch description
    int nshots = 0, kmax = 0, sg = 1, ngroups = 0;
    long long ntrace = 0;
    std::vector<FwiShotDev> shots_h;
    std::vector<int> nt_h, nr_h;
    FwiShotDev *shots_d = nullptr;
    float *srcf = nullptr;
    long long srcf_cap = 0;
    int *rrow = nullptr;
    int *rmap = nullptr;
    long long rmap_cap = 0;
    bool have_fwd = false, have_obs = false, have_res = false;
    // captured blocks of time steps
    int *kbase = nullptr;
    // parameter blocks of the step kernels that read them from device memory
    FwiFwdArgs fwd_args_h;
    FwiFwdArgs *fwd_args_d = nullptr;
    FwiAdjArgs adj_args_h;
    FwiAdjArgs *adj_args_d = nullptr;
    FwiRecArgs rec_args_h;
    FwiRecArgs *rec_args_d = nullptr;
    int graph_steps = 100;
    int fwd_kind = 0;  // 0: 2-D tile kernels, 1: z-marching kernels
    int rec_group = 8; // shots per workgroup of the 2-D tile reconstruction kernel
    int img_jump = 1, img_k0 = 0, img_k1 = INT_MAX; // imaging window of the adjoint pass
    // work lists of the 2-D tile kernels, [0] interior, [1] general (strips; imaging: group x tile)
    int *tl_fwd[2] = {nullptr, nullptr}, *tl_adj[2] = {nullptr, nullptr}, *tl_rec[2] = {nullptr, nullptr};
    int n_fwd[2] = {0, 0}, n_adj[2] = {0, 0}, n_rec[2] = {0, 0};
    int rec_lists_for = -1;  // rec_group the reconstruction lists were built for (-1: none)
    // fused backward step (tile kernels) on the padded grid's tiles: [0] fused
    // interior items (group, tile), group-major; [1] / [2] interior / general
    // adjoint strips (shot, tile) of the other tiles, shot-major; [3] their
    // reconstruction items (group, tile) with physical rows, group-major
    bool bwd_fused = true;
    int *tl_bwd[4] = {nullptr, nullptr, nullptr, nullptr};
    int n_bwd[4] = {0, 0, 0, 0};
    int bwd_lists_for = -1;  // rec_group the fused lists were built for (-1: none)
    std::vector<int> bwd_pre[4];
    // host copies of the shot tables last uploaded (set_shots skips an identical batch)
    std::vector<float> srcf_h;
    std::vector<int> rrow_h, rmap_h;  // list entries before group g ([0], [3]) or shot s ([1], [2])
    hipGraphExec_t gx_fwd = nullptr;
    std::vector<hipGraphExec_t> gx_bwd; // one captured block of backward steps per shot chunk
    int bwd_chunk = -1;                 // shots per backward pass (0: all in one, -1: auto_bwd_chunk)
    int bwd_lanes = 2;                  // backward chunks in flight (1 or 2; fwi_hip_set_bwd_lanes)
    // rows per wave of the fused backward step's general side when the
    // backward pass runs 4-row strips (4, or 8: the 8-row build's general
    // kernels on 8-row blocks; fwi_hip_set_bwd_general_rows)
    int gen_rows_bwd = 4;
    std::vector<int> adj_pre[2], rec_pre[2]; // list entries before shot s / before group g
    // timing: 0 off, 1 per launch (plain launches), 2 per pass (graphs as usual)
    int timing = 0;
    double t_ms[3] = {0, 0, 0};
    long long t_n[3] = {0, 0, 0};
    std::vector<hipEvent_t> ev;
};

static int set_err(fwi_hip_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) snprintf(c->err, sizeof(c->err), "%s", buf);
    snprintf(g_err, sizeof(g_err), "%s", buf);
    return code;
}

#define HIPCHK(c, call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return set_err((c), FWI_HIP_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                                      \
    } while (0)

static int dalloc(fwi_hip_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (c->d.hbm_budget && c->bytes + bytes > c->d.hbm_budget)
        return set_err(c, FWI_HIP_ENOMEM, "HBM budget exceeded: %zu + %zu > %zu bytes", c->bytes, bytes,
                       c->d.hbm_budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, FWI_HIP_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->bytes += bytes;
    c->allocs.push_back({*p, bytes});
    return FWI_HIP_OK;
}

static void dfree(fwi_hip_ctx *c, void *p)
{
    if (!p) return;
    auto it = std::find_if(c->allocs.begin(), c->allocs.end(), [p](const std::pair<void *, size_t> &a) {
        return a.first == p;
    });
    if (it != c->allocs.end()) {
        c->bytes -= it->second;
        c->allocs.erase(it);
    }
    (void)hipFree(p);
}

// --------------------------------------------------------------------------
// reference arithmetic on the host

// cpml_coe, fwi_2d.c:110-137: float storage, double where C promotes
extern "C" int fwi_hip_cpml_coe(float *a0, float *b0, int npml, float dx, float dt, float maxF, float maxV,
                                float coefR)
{
    if (!a0 || !b0 || npml < 1 || !(dx > 0) || !(dt > 0)) return set_err(nullptr, FWI_HIP_EINVAL, "cpml_coe: bad args");
    const float ndx = dx * 0.5;
    const int npower = 2; /* pml_base.dim at fwi_2d.c:257 */
    const double PI_REF = 3.141592654; /* the PI fwi_2d.c compiles with (fwi_sys.h) */
    float thickness_pml = npml * dx, PIf0 = PI_REF * maxF, slope = PIf0 / thickness_pml;
    float d0 = -(npower + 1) * maxV * log(coefR) / (2.0 * thickness_pml);
    for (int i = 0; i < 2 * npml + 1; i++) {
        float d = (1.0 - i * ndx / thickness_pml);
        d *= d;
        d *= d0;
        float alpha = slope * (i * ndx);
        b0[i] = exp(-(d + alpha) * dt);
        a0[i] = d / (d + alpha) * (b0[i] - 1.0);
    }
    return FWI_HIP_OK;
}

struct Coe {
    float c0x, c1x, ix, c0z, c1z, iz, iz840, iz24;
};

// fwi_fdcoe, fwi_2d.c:302-316
static Coe fd_coe(float dx, float dz)
{
    Coe c;
    c.ix = 1.0 / dx;
    c.iz = 1.0 / dz;
    c.c0x = 0.1129042e1 * c.ix;
    c.c1x = -0.4301412e-1 * c.ix;
    c.c0z = 0.1129042e1 * c.iz;
    c.c1z = -0.4301412e-1 * c.iz;
    c.iz840 = c.iz / 840.0;
    c.iz24 = c.iz / 24.0;
    return c;
}

// the six CPML rows used by the kernels: a0x b0x a1x b1x a1z b1z (mirror fwi_2d.c:275-283)
static void cpml_tables(const fwi_hip_grid &g, std::vector<float> &tab)
{
    const int T = 2 * g.npml + 1;
    std::vector<float> ax(T), bx(T), az(T), bz(T);
    fwi_hip_cpml_coe(ax.data(), bx.data(), g.npml, g.dx, g.dt, g.max_f, g.max_v, g.coef_r);
    fwi_hip_cpml_coe(az.data(), bz.data(), g.npml, g.dz, g.dt, g.max_f, g.max_v, g.coef_r);
    tab.assign(6 * T, 0.f);
    for (int i = 0; i < T; i++) {
        tab[0 * T + i] = ax[i];
        tab[1 * T + i] = bx[i];
        tab[2 * T + i] = ax[2 * g.npml - i];
        tab[3 * T + i] = bx[2 * g.npml - i];
        tab[4 * T + i] = az[2 * g.npml - i];
        tab[5 * T + i] = bz[2 * g.npml - i];
    }
}

static int check_grid(fwi_hip_ctx *c, const fwi_hip_grid &g)
{
    if (g.nx < 8 || g.nz < 8) return set_err(c, FWI_HIP_EINVAL, "grid %dx%d too small (need nx, nz >= 8)", g.nx, g.nz);
    if (g.npml < 2 || g.npml > 64) return set_err(c, FWI_HIP_EINVAL, "npml=%d: need 2..64 CPML cells", g.npml);
    if (!(g.dx > 0) || !(g.dz > 0) || !(g.dt > 0)) return set_err(c, FWI_HIP_EINVAL, "dx, dz, dt must be > 0");
    if (!(g.max_v > 0) || !(g.coef_r > 0) || !(g.coef_r < 1) || !(g.max_f >= 0))
        return set_err(c, FWI_HIP_EINVAL, "CPML meta: need max_v > 0, 0 < coef_r < 1, max_f >= 0");
    // one shot's padded field is addressed through a 32-bit buffer range
    const long long pnx = (long long)g.nx + 2 * g.npml, pnz = (long long)g.nz + g.npml;
    if (pnz * (FWI_TX * ((pnx + FWI_TX - 1) / FWI_TX)) * (long long)sizeof(float) >= FWI_BUF_OUT)
        return set_err(c, FWI_HIP_EINVAL, "grid %dx%d: one shot's padded field must stay below %d bytes", g.nx, g.nz,
                       FWI_BUF_OUT);
    return FWI_HIP_OK;
}

// --------------------------------------------------------------------------
extern "C" const char *fwi_hip_last_error(const fwi_hip_ctx *c) { return c ? c->err : g_err; }
static void drop_bwd(fwi_hip_ctx *c)
{
    for (hipGraphExec_t &x : c->gx_bwd)
        if (x) (void)hipGraphExecDestroy(x);
    c->gx_bwd.clear();
}

extern "C" size_t fwi_hip_bytes(const fwi_hip_ctx *c) { return c ? c->bytes : 0; }

extern "C" void fwi_hip_destroy(fwi_hip_ctx *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    drop_bwd(c);
    for (auto &a : c->allocs) (void)hipFree(a.first);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_fork2) (void)hipEventDestroy(c->ev_fork2);
    if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
    if (c->stB) (void)hipStreamDestroy(c->stB);
    if (c->stB2) (void)hipStreamDestroy(c->stB2);
    for (hipEvent_t e : {c->evB_fork, c->evB_join, c->evL_fork, c->evL_join, c->ev_adjd, c->ev_img[0], c->ev_img[1]})
        if (e) (void)hipEventDestroy(e);
    if (c->stc) (void)hipStreamSynchronize(c->stc);
    for (hipEvent_t e : c->blk_done) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->blk_copy) (void)hipEventDestroy(e);
    if (c->stc) (void)hipStreamDestroy(c->stc);
    delete c;
}

static int check_desc(fwi_hip_ctx *c, const fwi_hip_desc *desc)
{
    int rc = check_grid(c, desc->grid);
    if (!rc && (desc->max_shots < 1 || desc->max_nt < 1 || desc->max_nr < 0))
        rc = set_err(c, FWI_HIP_EINVAL, "max_shots, max_nt must be >= 1 and max_nr >= 0");
    if (!rc && desc->tile_z != 0 && desc->tile_z < 8) rc = set_err(c, FWI_HIP_EINVAL, "tile_z must be 0 or >= 8");
    if (!rc && (desc->shots_per_group < 0 || desc->shots_per_group > FWI_MAXSG))
        rc = set_err(c, FWI_HIP_EINVAL, "shots_per_group must be 0..%d", FWI_MAXSG);
    return rc;
}

// the device geometry of a descriptor (grid shapes, tiles, FD coefficients)
static FwiGeomDev make_geom(const fwi_hip_desc *desc)
{
    const fwi_hip_grid &gr = desc->grid;
    FwiGeomDev g;
    memset(&g, 0, sizeof(g));
    g.nx = gr.nx;
    g.nz = gr.nz;
    g.npml = gr.npml;
    g.pnx = gr.nx + 2 * gr.npml;
    g.pnz = gr.nz + gr.npml;
    g.ld = ((g.pnx + FWI_TX - 1) / FWI_TX) * FWI_TX;
    g.ntx = g.ld / FWI_TX;
    int tz = desc->tile_z;
    if (tz == 0) {
        const int nt_z = std::max(1, (g.pnz + 39) / 40);
        tz = std::max(8, (g.pnz + nt_z - 1) / nt_z);
    }
    g.tz = tz;
    g.ntz = (g.pnz + tz - 1) / tz;
    g.bntx = (g.nx + FWI_TX - 1) / FWI_TX;
    g.bntz = (g.nz + tz - 1) / tz;
    g.sl = 6 * g.nx + 6 * g.nz;
    g.ntz2 = (g.pnz + FWI_TZ2 - 1) / FWI_TZ2;
    Coe co = fd_coe(gr.dx, gr.dz);
    g.c0x = co.c0x; g.c1x = co.c1x; g.ix = co.ix;
    g.c0z = co.c0z; g.c1z = co.c1z; g.iz = co.iz;
    g.iz840 = co.iz840; g.iz24 = co.iz24;
    return g;
}

// strides of the per-shot field and CPML buffers of a geometry
static void set_strides(fwi_hip_ctx *c)
{
    const FwiGeomDev &g = c->g;
    const int n1 = g.npml + 1;
    c->fs = (long long)g.pnz * g.ld;
    c->xh_b = (long long)g.pnz * 2 * n1;
    c->xd_b = (long long)g.pnz * 2 * g.npml;
    c->zh_b = (long long)n1 * g.ld;
    c->zd_b = (long long)g.npml * g.ld;
}

// every device buffer a context allocates when it is created, with its size
// (the same table sizes the memory plan, fwi_hip_plan_memory)
static std::vector<std::pair<void **, size_t>> create_table(fwi_hip_ctx *c)
{
    const fwi_hip_desc *desc = &c->d;
    const fwi_hip_grid &gr = desc->grid;
    const FwiGeomDev &g = c->g;
    const long long S = desc->max_shots, NT = desc->max_nt, NR = desc->max_nr;
    const int maxgroups = (int)S;  // image accumulators: at most one group per shot, whatever the grouping
    // list capacity for the shortest strips (4 rows)
    const size_t strips = sizeof(int) * S * g.ntx * ((g.pnz + 3) / 4);
    const size_t items = sizeof(int) * S * g.bntx * ((g.nz + 3) / 4);
    return {
        {(void **)&c->ekp, sizeof(float) * c->fs},
        {(void **)&c->vel, sizeof(float) * gr.nx * gr.nz},
        {(void **)&c->P0, sizeof(float) * S * c->fs},
        {(void **)&c->P1, sizeof(float) * S * c->fs},
        {(void **)&c->M0, sizeof(float) * S * c->fs},
        {(void **)&c->M1, sizeof(float) * S * c->fs},
        {(void **)&c->xh, sizeof(float) * S * 2 * c->xh_b},
        {(void **)&c->xd, sizeof(float) * S * 2 * c->xd_b},
        {(void **)&c->zh, sizeof(float) * S * 2 * c->zh_b},
        {(void **)&c->zd, sizeof(float) * S * 2 * c->zd_b},
        {(void **)&c->tab, sizeof(float) * 6 * (2 * g.npml + 1)},
        {(void **)&c->syn_tm, sizeof(float) * S * NT * NR},
        {(void **)&c->syn, sizeof(float) * S * NT * NR},
        {(void **)&c->obs, sizeof(float) * S * NT * NR},
        {(void **)&c->res, sizeof(float) * S * NT * NR},
        {(void **)&c->res_tm, sizeof(float) * S * NT * NR},
        {(void **)&c->gacc, sizeof(double) * maxgroups * gr.nx * gr.nz},
        {(void **)&c->gk, sizeof(double) * gr.nx * gr.nz},
        {(void **)&c->grad, sizeof(double) * gr.nx * gr.nz},
        {(void **)&c->partial, sizeof(double) * c->npartial},
        {(void **)&c->shots_d, sizeof(FwiShotDev) * S},
        {(void **)&c->rrow, sizeof(int) * S * g.pnz},
        {(void **)&c->kbase, sizeof(int) * 64},
        {(void **)&c->fwd_args_d, sizeof(FwiFwdArgs)},
        {(void **)&c->tl_fwd[0], strips},
        {(void **)&c->tl_fwd[1], strips},
        {(void **)&c->fwd_work, strips + sizeof(int) * 64},
        {(void **)&c->tl_adj[0], strips},
        {(void **)&c->tl_adj[1], strips},
        {(void **)&c->tl_rec[0], items},
        {(void **)&c->tl_rec[1], items},
        {(void **)&c->tl_bwd[0], strips},
        {(void **)&c->tl_bwd[1], strips},
        {(void **)&c->tl_bwd[2], strips},
        {(void **)&c->tl_bwd[3], strips},
        {(void **)&c->adj_args_d, sizeof(FwiAdjArgs)},
        {(void **)&c->rec_args_d, sizeof(FwiRecArgs)},
        {(void **)&c->fwd_rc_d, sizeof(FwiFwdArgs)},
    };
}

extern "C" int fwi_hip_create(const fwi_hip_desc *desc, fwi_hip_ctx **out)
{
    if (!desc || !out) return set_err(nullptr, FWI_HIP_EINVAL, "fwi_hip_create: NULL argument");
    *out = nullptr;
    fwi_hip_ctx *c = new fwi_hip_ctx();
    c->d = *desc;
    const fwi_hip_grid &gr = desc->grid;
    int rc = check_desc(c, desc);
    if (rc) {
        snprintf(g_err, sizeof(g_err), "%s", c->err);
        delete c;
        return rc;
    }
    c->g = make_geom(desc);
    FwiGeomDev &g = c->g;
    c->sg = desc->shots_per_group ? desc->shots_per_group : std::min(FWI_MAXSG, desc->max_shots);

    if (desc->stream) {
        c->st = (hipStream_t)desc->stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            rc = set_err(c, FWI_HIP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
            delete c;
            return rc;
        }
        c->own_stream = true;
    }
    {
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess)
            c->lds_cu = lds;
        hipError_t e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stB, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stB2, hipStreamNonBlocking);
        for (hipEvent_t *ev : {&c->evB_fork, &c->evB_join, &c->evL_fork, &c->evL_join, &c->ev_adjd, &c->ev_img[0],
                               &c->ev_img[1]})
            if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            rc = set_err(c, FWI_HIP_EHIP, "second stream: %s", hipGetErrorString(e));
            fwi_hip_destroy(c);
            return rc;
        }
    }

    const long long S = desc->max_shots;
    set_strides(c);
    const auto plan = create_table(c);
    for (auto &a : plan) {
        rc = dalloc(c, a.first, a.second);
        if (rc) {
            char keep[512];
            snprintf(keep, sizeof(keep), "%s", c->err);
            fwi_hip_destroy(c);
            snprintf(g_err, sizeof(g_err), "%s", keep);
            return rc;
        }
    }
    // zero every field once so that the row padding stays zero for good
    long long nz_fields[] = {c->fs, S * c->fs, S * c->fs, S * c->fs, S * c->fs};
    float *fields[] = {c->ekp, c->P0, c->P1, c->M0, c->M1};
    for (int i = 0; i < 5; i++) {
        hipError_t e = fwi_launch_fill_f32(fields[i], nz_fields[i], 0.f, c->st);
        if (e != hipSuccess) {
            rc = set_err(c, FWI_HIP_EHIP, "fill: %s", hipGetErrorString(e));
            fwi_hip_destroy(c);
            return rc;
        }
    }
    std::vector<float> tab;
    cpml_tables(gr, tab);
    hipError_t e = hipMemcpyAsync(c->tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        rc = set_err(c, FWI_HIP_EHIP, "table upload: %s", hipGetErrorString(e));
        fwi_hip_destroy(c);
        return rc;
    }
    *out = c;
    return FWI_HIP_OK;
}

// Largest stable velocity of the leapfrog step p^{k+1} = 2 p^k - p^{k-1} +
// v^2 dt^2 L p^k: stable while v^2 dt^2 lambda_max(-L) <= 4 (von Neumann).
// -L splits into x and z parts whose largest eigenvalues add:
//   x: the staggered pair (c0, c1) of fd_coe at the Nyquist wavenumber,
//      4 (c0 - c1)^2 / dx^2 (the edge rows only lower it);
//   z: a mode bound to the free-surface rows (the one-sided 1/840 and 1/24
//      formulas of the top two v rows and the top vz row), 6.0195791 / dz^2,
//      the largest eigenvalue of the 1-D operator -D_vz D_v built from those
//      rows over the interior pair (it exceeds the interior 4 (c0 - c1)^2 by
//      a factor 1.0955 and does not depend on the depth of the grid).
static double stable_vmax(const fwi_hip_grid &g)
{
    const double c = 0.1129042e1 + 0.4301412e-1, lz_fs = 6.0195791;
    const double lx = 4.0 * c * c / ((double)g.dx * g.dx), lz = lz_fs / ((double)g.dz * g.dz);
    return 2.0 / ((double)g.dt * sqrt(lx + lz));
}

static int check_model(fwi_hip_ctx *c, const fwi_hip_grid &g, const float *vel, long long *cell)
{
    const long long n = (long long)g.nx * g.nz;
    long long imax = 0;
    for (long long i = 0; i < n; i++) {
        if (!(vel[i] > 0) || !isfinite(vel[i])) {
            if (cell) *cell = i;
            return set_err(c, FWI_HIP_EINVAL, "model: velocity must be finite and > 0 (cell %lld: %g)", i,
                           (double)vel[i]);
        }
        if (vel[i] > vel[imax]) imax = i;
    }
    const double vmax = stable_vmax(g);
    if (vel[imax] > vmax) {
        if (cell) *cell = imax;
        return set_err(c, FWI_HIP_EINVAL,
                       "model: velocity %g m/s at cell %lld is above the stability limit %g m/s of dt=%g, dx=%g, dz=%g",
                       (double)vel[imax], imax, vmax, (double)g.dt, (double)g.dx, (double)g.dz);
    }
    if (cell) *cell = -1;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_check_model(const fwi_hip_grid *g, const float *vel, long long *cell)
{
    if (!g || !vel || g->nx < 1 || g->nz < 1 || !(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0))
        return set_err(nullptr, FWI_HIP_EINVAL, "check_model: NULL argument or empty grid");
    return check_model(nullptr, *g, vel, cell);
}

// model_ext (fwi_2d.c:67-107) of kp = vel*vel*dt*dt into the pitched f-grid
extern "C" int fwi_hip_set_model(fwi_hip_ctx *c, const float *vel)
{
    if (!c || !vel) return set_err(c, FWI_HIP_EINVAL, "set_model: NULL argument");
    const FwiGeomDev &g = c->g;
    const float dt = c->d.grid.dt;
    const int nx = g.nx, nz = g.nz, npml = g.npml;
    // a model in device memory is checked and extended from a host copy
    std::vector<float> staged;
    hipPointerAttribute_t pa;
    if (hipPointerGetAttributes(&pa, vel) == hipSuccess && pa.type == hipMemoryTypeDevice) {
        staged.resize((size_t)nx * nz);
        HIPCHK(c, hipMemcpy(staged.data(), vel, sizeof(float) * staged.size(), hipMemcpyDefault));
        vel = staged.data();
    } else {
        (void)hipGetLastError();  // a plain host pointer is not an error
    }
    int rc = check_model(c, c->d.grid, vel, nullptr);
    if (rc) return rc;
    std::vector<float> kp((size_t)nx * nz), ekp((size_t)c->fs, 0.f);
    for (long long i = 0; i < (long long)nx * nz; i++) kp[i] = vel[i] * vel[i] * dt * dt;
    for (int z = 0; z < g.pnz; z++) {
        const int zs = std::min(z, nz - 1);
        for (int x = 0; x < g.pnx; x++) {
            const int xs = std::min(std::max(x - npml, 0), nx - 1);
            ekp[(size_t)z * g.ld + x] = kp[(size_t)zs * nx + xs];
        }
    }
    HIPCHK(c, hipMemcpyAsync(c->ekp, ekp.data(), sizeof(float) * ekp.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->vel, vel, sizeof(float) * nx * nz, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_model = true;
    return FWI_HIP_OK;
}

// The workgroup order of a one-launch step over the lists of its roles
// (count[role] workgroups each): block b runs on XCD b % 8, which takes a
// contiguous share of each list (neighbouring tiles share its L2, as
// xcd_remap arranges for the separate launches); within an XCD the roles come
// in `order`, the general ones first, since they live longest (1024x320 x 64
// shots, forward step: 58.5 us against 64.1 with the general workgroups spread
// evenly through the interior ones, and 59.4 on two streams).  Entry
// (role << 28) | index; -1 pads the XCDs to one length.
static std::vector<int> work_order(const std::vector<int> &count, const std::vector<int> &order)
{
    std::vector<std::vector<int>> seq(8);
    size_t len = 0;
    for (int x = 0; x < 8; x++) {
        for (int role : order) {
            const int n = count[role], per = (n + 7) / 8;
            const int i0 = std::min(x * per, n), i1 = std::min(i0 + per, n);
            for (int i = i0; i < i1; i++) seq[x].push_back((role << 28) | i);
        }
        len = std::max(len, seq[x].size());
    }
    std::vector<int> w(8 * len, -1);
    for (int x = 0; x < 8; x++)
        for (size_t j = 0; j < seq[x].size(); j++) w[j * 8 + x] = seq[x][j];
    return w;
}

// strip lists of the forward (strip_fwd rows) and adjoint (strip_bwd rows)
// tile kernels, logical strip (s nsz + sz) ntx + tx, for the context's shots,
// and the workgroup order of the one-launch forward step; the backward pass
// builds its reconstruction and fused lists itself
static int build_strip_lists(fwi_hip_ctx *c)
{
    const FwiGeomDev &g = c->g;
    const int nshots = c->nshots;
    std::vector<int> lf[2], la[2];
    for (int q = 0; q < 2; q++) c->adj_pre[q].assign(nshots + 1, 0);
    const int fr = c->strip_fwd, br = c->strip_bwd;
    const int nszf = (g.pnz + fr - 1) / fr, nszb = (g.pnz + br - 1) / br;
    for (int s = 0; s < nshots; s++) {
        for (int q = 0; q < 2; q++) c->adj_pre[q][s] = (int)la[q].size();
        for (int sz = 0; sz < nszf; sz++)
            for (int tx = 0; tx < g.ntx; tx++)
                lf[fwi_fwd_strip_interior(g, c->shots_h[s], tx * FWI_TX, sz * fr, fr) ? 0 : 1].push_back(
                    (s * nszf + sz) * g.ntx + tx);
        for (int sz = 0; sz < nszb; sz++)
            for (int tx = 0; tx < g.ntx; tx++)
                la[fwi_adj_strip_interior(g, c->shots_h[s], tx * FWI_TX, sz * br, br) ? 0 : 1].push_back(
                    (s * nszb + sz) * g.ntx + tx);
    }
    for (int q = 0; q < 2; q++) c->adj_pre[q][nshots] = (int)la[q].size();
    for (int q = 0; q < 2; q++) {
        c->n_fwd[q] = (int)lf[q].size();
        c->n_adj[q] = (int)la[q].size();
        if (c->n_fwd[q])
            HIPCHK(c, hipMemcpyAsync(c->tl_fwd[q], lf[q].data(), sizeof(int) * lf[q].size(), hipMemcpyHostToDevice,
                                     c->st));
        if (c->n_adj[q])
            HIPCHK(c, hipMemcpyAsync(c->tl_adj[q], la[q].data(), sizeof(int) * la[q].size(), hipMemcpyHostToDevice,
                                     c->st));
    }
    const std::vector<int> work = work_order({(c->n_fwd[0] + 3) / 4, (c->n_fwd[1] + 3) / 4}, {1, 0});
    c->n_fwd_work = (int)work.size();
    if (!work.empty())
        HIPCHK(c, hipMemcpyAsync(c->fwd_work, work.data(), sizeof(int) * work.size(), hipMemcpyHostToDevice, c->st));
    c->rec_lists_for = -1;
    c->bwd_lists_for = -1;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_shots(fwi_hip_ctx *c, int nshots, const fwi_hip_shot *shots)
{
    if (!c || !shots) return set_err(c, FWI_HIP_EINVAL, "set_shots: NULL argument");
    if (nshots < 1 || nshots > c->d.max_shots)
        return set_err(c, FWI_HIP_EINVAL, "set_shots: nshots=%d outside 1..%d", nshots, c->d.max_shots);
    const FwiGeomDev &g = c->g;
    std::vector<FwiShotDev> sd(nshots);
    std::vector<int> rrow((size_t)nshots * g.pnz, -1);
    std::vector<int> rmap;
    std::vector<float> srcf;
    const float iarea = 1.0 / (c->d.grid.dx * c->d.grid.dz); /* fwi_2d.c:216 */
    long long toff = 0, soff = 0;
    int slots = 0, kmax = 0;
    c->nt_h.assign(nshots, 0);
    c->nr_h.assign(nshots, 0);
    for (int s = 0; s < nshots; s++) {
        const fwi_hip_shot &sh = shots[s];
        if (sh.nt < 1 || sh.nt > c->d.max_nt)
            return set_err(c, FWI_HIP_EINVAL, "shot %d: nt=%d outside 1..%d", s, sh.nt, c->d.max_nt);
        if (sh.nr < 0 || sh.nr > c->d.max_nr)
            return set_err(c, FWI_HIP_EINVAL, "shot %d: nr=%d outside 0..%d", s, sh.nr, c->d.max_nr);
        if (!sh.wav || (sh.nr > 0 && (!sh.irx || !sh.irz)))
            return set_err(c, FWI_HIP_EINVAL, "shot %d: NULL wavelet or receiver arrays", s);
        if (sh.isx < 0 || sh.isx >= g.nx || sh.isz < 0 || sh.isz >= g.nz)
            return set_err(c, FWI_HIP_EINVAL, "shot %d: source cell (%d,%d) outside the %dx%d grid", s, sh.isx,
                           sh.isz, g.nx, g.nz);
        FwiShotDev &d = sd[s];
        d.nt = sh.nt;
        d.isx = sh.isx + g.npml;
        d.isz = sh.isz;
        d.nr = sh.nr;
        d.src_off = (long long)srcf.size();
        for (int k = 0; k < sh.nt; k++) srcf.push_back(sh.wav[k] * iarea);
        d.trace_off = toff;
        toff += (long long)sh.nr * sh.nt;
        d.strip_off = soff;
        soff += (long long)sh.nt * g.sl;
        d.rrow_off = s * g.pnz;
        d.rz_lo = g.pnz;
        d.rz_hi = -1;
        std::map<int, int> rowslot;
        for (int r = 0; r < sh.nr; r++) {
            const int x = sh.irx[r], z = sh.irz[r];
            if (x < 0 || x >= g.nx || z < 0 || z >= g.nz)
                return set_err(c, FWI_HIP_EINVAL, "shot %d receiver %d: cell (%d,%d) outside the grid", s, r, x, z);
            auto it = rowslot.find(z);
            int slot;
            if (it == rowslot.end()) {
                slot = slots++;
                rowslot[z] = slot;
                rrow[(size_t)s * g.pnz + z] = slot;
                rmap.resize((size_t)slots * g.ld, -1);
            } else {
                slot = it->second;
            }
            int &cell = rmap[(size_t)slot * g.ld + x + g.npml];
            if (cell >= 0)
                return set_err(c, FWI_HIP_EINVAL, "shot %d: receivers %d and %d share cell (%d,%d)", s, cell, r, x, z);
            cell = r;
            d.rz_lo = std::min(d.rz_lo, z);
            d.rz_hi = std::max(d.rz_hi, z);
        }
        c->nt_h[s] = sh.nt;
        c->nr_h[s] = sh.nr;
        kmax = std::max(kmax, sh.nt);
    }
    if (toff > (long long)c->d.max_shots * c->d.max_nt * c->d.max_nr)
        return set_err(c, FWI_HIP_EINVAL, "set_shots: traces exceed the context capacity");
    if (rmap.empty()) rmap.assign(g.ld, -1);
    // the same shots as the device holds (a driver that sets its shots before
    // every gradient): keep the tables, strip lists and captured graphs
    if (nshots == c->nshots && kmax == c->kmax && srcf == c->srcf_h && rrow == c->rrow_h && rmap == c->rmap_h &&
        memcmp(sd.data(), c->shots_h.data(), sizeof(FwiShotDev) * nshots) == 0) {
        c->have_fwd = c->have_obs = c->have_res = false;
        return FWI_HIP_OK;
    }
    c->srcf_h.clear();  // no match until the new tables are on the device
    // (re)allocate the per-batch tables
    if ((long long)srcf.size() > c->srcf_cap) {
        dfree(c, c->srcf);
        c->srcf = nullptr;
        c->srcf_cap = 0;
        int rc = dalloc(c, (void **)&c->srcf, sizeof(float) * srcf.size());
        if (rc) return rc;
        c->srcf_cap = (long long)srcf.size();
    }
    if ((long long)rmap.size() > c->rmap_cap) {
        dfree(c, c->rmap);
        c->rmap = nullptr;
        c->rmap_cap = 0;
        int rc = dalloc(c, (void **)&c->rmap, sizeof(int) * rmap.size());
        if (rc) return rc;
        c->rmap_cap = (long long)rmap.size();
    }
    HIPCHK(c, hipMemcpyAsync(c->srcf, srcf.data(), sizeof(float) * srcf.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->rmap, rmap.data(), sizeof(int) * rmap.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->rrow, rrow.data(), sizeof(int) * rrow.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->shots_d, sd.data(), sizeof(FwiShotDev) * nshots, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);
    c->shots_h = sd;
    c->srcf_h = std::move(srcf);
    c->rrow_h = std::move(rrow);
    c->rmap_h = std::move(rmap);
    c->nshots = nshots;
    c->kmax = kmax;
    c->ntrace = toff;
    c->ngroups = (nshots + c->sg - 1) / c->sg;
    c->have_fwd = c->have_obs = c->have_res = false;
    return build_strip_lists(c);
}

// dynamic LDS per interior workgroup so that int_wgs of them, and no more,
// fit on a CU beside one general workgroup (0 when not capped)
static int interior_lds(const fwi_hip_ctx *c, int fr, int n)
{
    if (!c->concurrent || n <= 0 || c->lds_cu <= 0) return 0;
    // the least LDS per interior workgroup that keeps n + 1 of them off a CU,
    // so that the rest holds as many general workgroups as it can
    const int total = (c->lds_cu / (n + 1) + 256) & ~255;
    if (total * n + fwi_lds_gen(fr) > c->lds_cu || total * (n + 1) <= c->lds_cu || total <= fwi_lds_tile(fr))
        return 0;
    return total - fwi_lds_tile(fr);
}

// streams and occupancy cap of a pass (0 forward, 1 backward)
static FwiPair pair_of(const fwi_hip_ctx *c, int pass)
{
    const int fr = pass ? c->strip_bwd : c->strip_fwd;
    FwiPair p;
    p.st = c->st;
    p.st2 = c->concurrent ? c->st2 : nullptr;
    p.fork = c->ev_fork;
    p.join = c->ev_join;
    p.lds_int = interior_lds(c, fr, pass ? c->int_wgs_bwd : c->int_wgs_fwd);
    return p;
}

// the tile kernels of the strip height of the pass (fwi_hip_set_pass_shape)
// whether the fused backward lists are built with 8-row blocks on the general
// side under 4-row interior tiles (the one-launch step needs one strip height)
static bool tall_general(const fwi_hip_ctx *c)
{
    return c->strip_bwd == 4 && c->gen_rows_bwd == 8 && !c->bwd_mixed;
}

static hipError_t fwd_tile(const fwi_hip_ctx *c, const FwiFwdArgs *a, const FwiFwdArgs *a_dev, const int *const *tl,
                           const int *ntl, const int *kbase, int koff, const FwiPair &p, int mode)
{
    if (c->fwd_mixed && tl == c->tl_fwd)
        return c->strip_fwd == 4
                   ? fwi_launch_fwd_mixed_s4(a, a_dev, c->fwd_work, c->n_fwd_work, tl, ntl, kbase, koff, p.st, mode)
                   : fwi_launch_fwd_mixed(a, a_dev, c->fwd_work, c->n_fwd_work, tl, ntl, kbase, koff, p.st, mode);
    return c->strip_fwd == 4 ? fwi_launch_fwd_tile_s4(a, a_dev, tl, ntl, kbase, koff, p, mode)
                             : fwi_launch_fwd_tile(a, a_dev, tl, ntl, kbase, koff, p, mode);
}
static hipError_t adj_tile(const fwi_hip_ctx *c, const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const int *const *tl,
                           const int *ntl, const int *kbase, int koff, const FwiPair &p)
{
    return c->strip_bwd == 4 ? fwi_launch_adj_tile_s4(a, a_dev, tl, ntl, kbase, koff, p)
                             : fwi_launch_adj_tile(a, a_dev, tl, ntl, kbase, koff, p);
}
static hipError_t rec_tile(const fwi_hip_ctx *c, const FwiRecArgs *a, const FwiRecArgs *a_dev, const int *const *tl,
                           const int *ntl, const int *kbase, int koff, int gsize, const FwiPair &p)
{
    return c->strip_bwd == 4 ? fwi_launch_rec_tile_s4(a, a_dev, tl, ntl, kbase, koff, gsize, p)
                             : fwi_launch_rec_tile(a, a_dev, tl, ntl, kbase, koff, gsize, p);
}
static hipError_t bwd_tile(const fwi_hip_ctx *c, const FwiAdjArgs *a, const FwiAdjArgs *a_dev, const FwiRecArgs *r_dev,
                           const int *const *tl, const int *ntl, const int *kbase, int koff, int gsize,
                           const FwiPair &p)
{
    if (tall_general(c)) {
        // the interior side of the 4-row build beside the general side of
        // the 8-row build, forked and joined as fwi_launch_bwd_tile does
        const bool has_int = ntl[0] > 0, has_gen = ntl[1] > 0 || ntl[2] > 0 || ntl[3] > 0;
        if (p.st2 && has_int && has_gen) {
            hipError_t e = hipEventRecord(p.fork, p.st);
            if (e == hipSuccess) e = hipStreamWaitEvent(p.st2, p.fork, 0);
            if (e == hipSuccess) e = fwi_launch_bwd_gen(a, a_dev, r_dev, tl, ntl, kbase, koff, gsize, p.st2);
            if (e == hipSuccess)
                e = fwi_launch_bwd_int_s4(a, a_dev, r_dev, tl[0], ntl[0], kbase, koff, gsize, p.lds_int, p.st);
            if (e == hipSuccess) e = hipEventRecord(p.join, p.st2);
            if (e == hipSuccess) e = hipStreamWaitEvent(p.st, p.join, 0);
            return e;
        }
        hipError_t e = fwi_launch_bwd_int_s4(a, a_dev, r_dev, tl[0], ntl[0], kbase, koff, gsize, p.lds_int, p.st);
        if (e == hipSuccess && has_gen) e = fwi_launch_bwd_gen(a, a_dev, r_dev, tl, ntl, kbase, koff, gsize, p.st);
        return e;
    }
    return c->strip_bwd == 4 ? fwi_launch_bwd_tile_s4(a, a_dev, r_dev, tl, ntl, kbase, koff, gsize, p)
                             : fwi_launch_bwd_tile(a, a_dev, r_dev, tl, ntl, kbase, koff, gsize, p);
}

static FwiBatchDev batch_dev(fwi_hip_ctx *c)
{
    FwiBatchDev b;
    memset(&b, 0, sizeof(b));
    b.nshots = c->nshots;
    b.sg = c->sg;
    b.ngroups = c->ngroups;
    b.shots = c->shots_d;
    b.fs = c->fs;
    b.xh_b = c->xh_b;
    b.xd_b = c->xd_b;
    b.zh_b = c->zh_b;
    b.zd_b = c->zd_b;
    b.xh_s = 2 * c->xh_b;
    b.xd_s = 2 * c->xd_b;
    b.zh_s = 2 * c->zh_b;
    b.zd_s = 2 * c->zd_b;
    b.ekp = c->ekp;
    b.tab = c->tab;
    b.srcf = c->srcf;
    b.rrow = c->rrow;
    b.rmap = c->rmap;
    return b;
}

// boundary frames of the boundary-saving reconstruction (max_shots x max_nt steps)
static int ensure_frames(fwi_hip_ctx *c)
{
    if (c->pstrip) return FWI_HIP_OK;
    const size_t n = sizeof(float) * (size_t)c->d.max_shots * c->d.max_nt * c->g.sl;
    int rc = dalloc(c, (void **)&c->pstrip, n);
    if (!rc) rc = dalloc(c, (void **)&c->xstrip, n);
    if (rc) {
        dfree(c, c->pstrip);
        dfree(c, c->xstrip);
        c->pstrip = c->xstrip = nullptr;
    }
    return rc;
}

static long long cpml_floats(const fwi_hip_ctx *c, long long S)
{
    return S * 2 * (c->xh_b + c->xd_b + c->zh_b + c->zd_b);
}

static int ensure_cpml2(fwi_hip_ctx *c);
static void free_ckpt(fwi_hip_ctx *c);

// snapshot, scale and segment buffers for the current shots and nt
static int ensure_ckpt(fwi_hip_ctx *c)
{
    const long long S = c->nshots, K = c->ckpt_every, nseg = (c->kmax + K - 1) / K;
    const long long nb = (long long)c->g.nx * c->g.nz;
    if (c->ck_store && c->ck_K == K && nseg <= c->ck_nseg && S <= c->ck_S) return FWI_HIP_OK;
    free_ckpt(c);
    int rc = ensure_cpml2(c);
    if (rc) return rc;
    const size_t es = c->ckpt_storage ? 2 : 4;
    if ((rc = dalloc(c, &c->ck_store, es * nseg * 2 * S * c->fs)) ||
        (rc = dalloc(c, (void **)&c->ck_scale, sizeof(unsigned) * nseg * 2 * S)) ||
        (rc = dalloc(c, (void **)&c->ck_cpml, sizeof(float) * nseg * cpml_floats(c, S))) ||
        (rc = dalloc(c, (void **)&c->ck_xseg, sizeof(float) * K * S * nb)))
        return rc;
    c->ck_nseg = nseg;
    c->ck_S = S;
    c->ck_K = K;
    return FWI_HIP_OK;
}

// The forward state before step k (p^k, p^{k-1} in P0/P1 by parity, and the
// CPML memory of the forward pass, or of the recompute when `recompute`) into
// snapshot slot j.
static int ck_save_at(fwi_hip_ctx *c, long long j, long long k, bool recompute)
{
    const long long S = c->nshots, n = S * c->fs;
    float *cp = c->ck_cpml + j * cpml_floats(c, c->ck_S);
    const float *xh = recompute ? c->xh2 : c->xh, *xd = recompute ? c->xd2 : c->xd;
    const float *zh = recompute ? c->zh2 : c->zh, *zd = recompute ? c->zd2 : c->zd;
    HIPCHK(c, hipMemcpyAsync(cp, xh, sizeof(float) * S * 2 * c->xh_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->xh_b;
    HIPCHK(c, hipMemcpyAsync(cp, xd, sizeof(float) * S * 2 * c->xd_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->xd_b;
    HIPCHK(c, hipMemcpyAsync(cp, zh, sizeof(float) * S * 2 * c->zh_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->zh_b;
    HIPCHK(c, hipMemcpyAsync(cp, zd, sizeof(float) * S * 2 * c->zd_b, hipMemcpyDeviceToDevice, c->st));
    if (c->ckpt_storage == 0) {
        float *dst = (float *)c->ck_store + j * 2 * c->ck_S * c->fs;
        HIPCHK(c, hipMemcpyAsync(dst, c->P0, sizeof(float) * n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(dst + n, c->P1, sizeof(float) * n, hipMemcpyDeviceToDevice, c->st));
        return FWI_HIP_OK;
    }
    // 16 bit: p^k and the step difference p^k - p^{k-1}, each with its own
    // per-shot scale.  Rounding the two time levels independently would leave
    // their difference (the time derivative, a few per cent of p at a few
    // tens of steps per period) with the rounding of p itself: tens of times
    // the storage precision after the restart.
    float *pk = (k & 1) ? c->P1 : c->P0, *pm = (k & 1) ? c->P0 : c->P1;  // p^k, p^{k-1}
    unsigned *bits = c->ck_scale + j * 2 * c->ck_S;
    unsigned short *dst = (unsigned short *)c->ck_store + j * 2 * c->ck_S * c->fs;
    const int fp16 = c->ckpt_storage == 1;
    HIPCHK(c, fwi_launch_fill_f32((float *)bits, 2 * S, 0.f, c->st));
    HIPCHK(c, fwi_launch_absmax(pk, nullptr, c->fs, (int)S, bits, c->st));
    HIPCHK(c, fwi_launch_absmax(pk, pm, c->fs, (int)S, bits + S, c->st));
    HIPCHK(c, fwi_launch_pack16(pk, nullptr, c->fs, (int)S, bits, dst, fp16, c->st));
    HIPCHK(c, fwi_launch_pack16(pk, pm, c->fs, (int)S, bits + S, dst + n, fp16, c->st));
    return FWI_HIP_OK;
}

// snapshot slot j (taken before step k) back into P0/P1 and the recompute's CPML set
static int ck_restore_at(fwi_hip_ctx *c, long long j, long long k)
{
    const long long S = c->nshots, n = S * c->fs;
    const float *cp = c->ck_cpml + j * cpml_floats(c, c->ck_S);
    HIPCHK(c, hipMemcpyAsync(c->xh2, cp, sizeof(float) * S * 2 * c->xh_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->xh_b;
    HIPCHK(c, hipMemcpyAsync(c->xd2, cp, sizeof(float) * S * 2 * c->xd_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->xd_b;
    HIPCHK(c, hipMemcpyAsync(c->zh2, cp, sizeof(float) * S * 2 * c->zh_b, hipMemcpyDeviceToDevice, c->st));
    cp += S * 2 * c->zh_b;
    HIPCHK(c, hipMemcpyAsync(c->zd2, cp, sizeof(float) * S * 2 * c->zd_b, hipMemcpyDeviceToDevice, c->st));
    if (c->ckpt_storage == 0) {
        const float *src = (const float *)c->ck_store + j * 2 * c->ck_S * c->fs;
        HIPCHK(c, hipMemcpyAsync(c->P0, src, sizeof(float) * n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(c->P1, src + n, sizeof(float) * n, hipMemcpyDeviceToDevice, c->st));
        return FWI_HIP_OK;
    }
    float *pk = (k & 1) ? c->P1 : c->P0, *pm = (k & 1) ? c->P0 : c->P1;
    const unsigned *bits = c->ck_scale + j * 2 * c->ck_S;
    const unsigned short *src = (const unsigned short *)c->ck_store + j * 2 * c->ck_S * c->fs;
    const int fp16 = c->ckpt_storage == 1;
    HIPCHK(c, fwi_launch_unpack16(src, c->fs, (int)S, bits, nullptr, pk, fp16, c->st));
    HIPCHK(c, fwi_launch_unpack16(src + n, c->fs, (int)S, bits + S, pk, pm, fp16, c->st));  // p^{k-1} = p^k - d
    return FWI_HIP_OK;
}

// uniform snapshots: slot j holds the state before step j K
static int ck_save(fwi_hip_ctx *c, long long j) { return ck_save_at(c, j, j * c->ckpt_every, false); }
static int ck_restore(fwi_hip_ctx *c, long long j) { return ck_restore_at(c, j, j * c->ckpt_every); }

// ---- binomial checkpointing: buffers and the memory plan
#define FWI_RV_MAX_SNAPS 128

static int ensure_cpml2(fwi_hip_ctx *c)
{
    if (c->xh2) return FWI_HIP_OK;
    int rc;
    if ((rc = dalloc(c, (void **)&c->xh2, sizeof(float) * c->d.max_shots * 2 * c->xh_b)) ||
        (rc = dalloc(c, (void **)&c->xd2, sizeof(float) * c->d.max_shots * 2 * c->xd_b)) ||
        (rc = dalloc(c, (void **)&c->zh2, sizeof(float) * c->d.max_shots * 2 * c->zh_b)) ||
        (rc = dalloc(c, (void **)&c->zd2, sizeof(float) * c->d.max_shots * 2 * c->zd_b)))
        return rc;
    return FWI_HIP_OK;
}

static void free_ckpt(fwi_hip_ctx *c)
{
    dfree(c, c->ck_store);
    dfree(c, c->ck_scale);
    dfree(c, c->ck_cpml);
    dfree(c, c->ck_xseg);
    c->ck_store = nullptr;
    c->ck_scale = nullptr;
    c->ck_cpml = c->ck_xseg = nullptr;
    c->ck_nseg = c->ck_S = c->ck_K = 0;
}

static void free_frames(fwi_hip_ctx *c)
{
    dfree(c, c->pstrip);
    dfree(c, c->xstrip);
    dfree(c, c->xg);
    c->pstrip = c->xstrip = c->xg = nullptr;
}

// bytes of each backward method's buffers, at the context's capacity
static size_t frames_bytes(const fwi_hip_ctx *c)
{
    return 2 * sizeof(float) * (size_t)c->d.max_shots * (size_t)c->d.max_nt * (size_t)c->g.sl +
           (c->bwd_mixed ? 2 * sizeof(float) * (size_t)c->d.max_shots * (size_t)c->g.nx * (size_t)c->g.nz : 0);
}
static size_t rv_slot_bytes(const fwi_hip_ctx *c, int storage)
{
    const size_t S = c->d.max_shots, es = storage ? 2 : 4;
    return es * 2 * S * (size_t)c->fs + sizeof(float) * (size_t)cpml_floats(c, S) +
           (storage ? sizeof(unsigned) * 2 * S : 0);
}
// the recompute's CPML set and one step of X
static size_t rv_fixed_bytes(const fwi_hip_ctx *c)
{
    const size_t S = c->d.max_shots;
    return sizeof(float) * (size_t)cpml_floats(c, S) + sizeof(float) * S * (size_t)c->g.nx * c->g.nz;
}

// The backward method within the HBM budget beside `base` bytes: the boundary
// frames when they fit (or without a budget), else binomial checkpoints with
// as many slots as fit (at most FWI_RV_MAX_SNAPS, plus the slot of the state
// before the last step).
static int plan_core(fwi_hip_ctx *c, size_t base, int storage, int *method, int *snaps)
{
    const size_t B = c->d.hbm_budget, fr = frames_bytes(c);
    *method = 0;
    *snaps = 0;
    if (!B || base + fr <= B) return FWI_HIP_OK;
    const size_t fixed = rv_fixed_bytes(c), slot = rv_slot_bytes(c, storage);
    const long long n = (B > base + fixed) ? (long long)((B - base - fixed) / slot) : 0;
    const long long sn = std::min<long long>(n - 1, FWI_RV_MAX_SNAPS);
    if (sn < 2)
        return set_err(c, FWI_HIP_ENOMEM,
                       "HBM budget %zu bytes: beside the context's %zu, neither the boundary frames (%zu) nor 3 "
                       "snapshots (%zu each, + %zu) fit",
                       B, base, fr, slot, fixed);
    *method = 2;
    *snaps = (int)sn;
    return FWI_HIP_OK;
}

// device bytes held beyond what the backward method decides
static size_t base_bytes(const fwi_hip_ctx *c)
{
    size_t held = c->bytes;
    for (const auto &a : c->allocs)
        if (a.first == c->pstrip || a.first == c->xstrip || a.first == c->xg || a.first == c->ck_store || a.first == c->ck_scale ||
            a.first == c->ck_cpml || a.first == c->ck_xseg || a.first == c->xh2 || a.first == c->xd2 ||
            a.first == c->zh2 || a.first == c->zd2)
            held -= a.second;
    return held;
}

// slots (and the one-step X buffer) of the binomial schedule, at capacity
static int ensure_slots(fwi_hip_ctx *c, long long nslots)
{
    const long long S = c->d.max_shots, nb = (long long)c->g.nx * c->g.nz;
    if (c->ck_store && c->ck_K == 1 && nslots <= c->ck_nseg && S <= c->ck_S) return FWI_HIP_OK;
    free_ckpt(c);
    int rc = ensure_cpml2(c);
    if (rc) return rc;
    const size_t es = c->ckpt_storage ? 2 : 4;
    if ((rc = dalloc(c, &c->ck_store, es * nslots * 2 * S * c->fs)) ||
        (rc = dalloc(c, (void **)&c->ck_scale, sizeof(unsigned) * nslots * 2 * S)) ||
        (rc = dalloc(c, (void **)&c->ck_cpml, sizeof(float) * nslots * cpml_floats(c, S))) ||
        (rc = dalloc(c, (void **)&c->ck_xseg, sizeof(float) * S * nb)))
        return rc;
    c->ck_nseg = nslots;
    c->ck_S = S;
    c->ck_K = 1;
    return FWI_HIP_OK;
}

static int zero_cpml(fwi_hip_ctx *c)
{
    const long long S = c->nshots;
    HIPCHK(c, fwi_launch_fill_f32(c->xh, S * 2 * c->xh_b, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f32(c->xd, S * 2 * c->xd_b, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f32(c->zh, S * 2 * c->zh_b, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f32(c->zd, S * 2 * c->zd_b, 0.f, c->st));
    return FWI_HIP_OK;
}

// timing helpers: an event pair around a launch (timing 1)
static int tick(fwi_hip_ctx *c, hipEvent_t *e)
{
    *e = nullptr;
    if (c->timing != 1) return FWI_HIP_OK;
    HIPCHK(c, hipEventCreate(e));
    c->ev.push_back(*e);
    HIPCHK(c, hipEventRecord(*e, c->st));
    return FWI_HIP_OK;
}

// an event pair around a whole pass (timing 2): t_ms[which] += its time,
// t_n[which] += its time steps
// the copy stream and n event pairs
static int ensure_copy(fwi_hip_ctx *c, int n)
{
    if (!c->stc) HIPCHK(c, hipStreamCreateWithFlags(&c->stc, hipStreamNonBlocking));
    while ((int)c->blk_done.size() < n) {
        hipEvent_t a, b;
        HIPCHK(c, hipEventCreateWithFlags(&a, hipEventDisableTiming));
        c->blk_done.push_back(a);
        HIPCHK(c, hipEventCreateWithFlags(&b, hipEventDisableTiming));
        c->blk_copy.push_back(b);
    }
    return FWI_HIP_OK;
}

// rows [k0, k1) of every shot's time-major traces between host and device
// (each shot's rows are contiguous at its trace offset; one strided copy when
// the shots are packed alike)
static hipError_t copy_rows(fwi_hip_ctx *c, float *host, float *dev, int k0, int k1, bool to_host, hipStream_t s)
{
    const hipMemcpyKind kind = hipMemcpyDefault;  // the caller's buffer may be host or device memory
    const FwiShotDev &d0 = c->shots_h[0];
    bool alike = d0.nr > 0;
    for (int i = 0; i < c->nshots && alike; i++) {
        const FwiShotDev &d = c->shots_h[i];
        alike = d.nt == d0.nt && d.nr == d0.nr && d.trace_off == (long long)i * d0.nt * d0.nr;
    }
    if (alike) {
        const int k1c = std::min(k1, d0.nt);
        if (k0 >= k1c) return hipSuccess;
        const size_t pitch = sizeof(float) * (size_t)d0.nt * d0.nr, off = (size_t)k0 * d0.nr;
        const size_t width = sizeof(float) * (size_t)(k1c - k0) * d0.nr;
        return to_host ? hipMemcpy2DAsync(host + off, pitch, dev + off, pitch, width, c->nshots, kind, s)
                       : hipMemcpy2DAsync(dev + off, pitch, host + off, pitch, width, c->nshots, kind, s);
    }
    for (int i = 0; i < c->nshots; i++) {
        const FwiShotDev &d = c->shots_h[i];
        const int k1c = std::min(k1, d.nt);
        if (d.nr == 0 || k0 >= k1c) continue;
        const size_t off = (size_t)d.trace_off + (size_t)k0 * d.nr, n = sizeof(float) * (size_t)(k1c - k0) * d.nr;
        hipError_t e = to_host ? hipMemcpyAsync(host + off, dev + off, n, kind, s)
                               : hipMemcpyAsync(dev + off, host + off, n, kind, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int pass_begin(fwi_hip_ctx *c, hipEvent_t *e)
{
    *e = nullptr;
    if (c->timing != 2) return FWI_HIP_OK;
    HIPCHK(c, hipEventCreate(e));
    HIPCHK(c, hipEventRecord(*e, c->st));
    return FWI_HIP_OK;
}
static int pass_end(fwi_hip_ctx *c, int which, hipEvent_t e0, long long steps)
{
    if (!e0) return FWI_HIP_OK;
    hipEvent_t e1 = nullptr;
    float ms = 0.f;
    hipError_t e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e1, c->st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return set_err(c, FWI_HIP_EHIP, "pass timing: %s", hipGetErrorString(e));
    c->t_ms[which] += ms;
    c->t_n[which] += steps;
    return FWI_HIP_OK;
}

static int collect_timing(fwi_hip_ctx *c, std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> &marks)
{
    if (marks.empty()) return FWI_HIP_OK;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto &m : marks) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, m.second.first, m.second.second));
        c->t_ms[m.first] += ms;
        c->t_n[m.first] += 1;
    }
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    c->ev.clear();
    marks.clear();
    return FWI_HIP_OK;
}

// Captures `build` (the launches of graph_steps time steps plus the counter
// advance) into an executable graph.
template <class F>
static int capture(fwi_hip_ctx *c, hipGraphExec_t *out, F build)
{
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    hipError_t e = build();
    hipError_t e2 = hipStreamEndCapture(c->st, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_err(c, FWI_HIP_EHIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *out = nullptr;
        return set_err(c, FWI_HIP_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_kernels(fwi_hip_ctx *c, int fwd_kind)
{
    if (!c || fwd_kind < 0 || fwd_kind > 1) return set_err(c, FWI_HIP_EINVAL, "set_kernels: bad argument");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);
    c->fwd_kind = fwd_kind;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_rec_group(fwi_hip_ctx *c, int shots)
{
    if (!c || shots < 1) return set_err(c, FWI_HIP_EINVAL, "set_rec_group: need shots >= 1");
    drop_bwd(c);
    c->rec_group = shots;
    c->rec_lists_for = -1;
    c->bwd_lists_for = -1;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_concurrent(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_concurrent: NULL context");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);
    c->concurrent = on != 0;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_pass_shape(fwi_hip_ctx *c, int pass, int rows, int interior_wgs)
{
    if (!c || (pass != 0 && pass != 1) || (rows != 4 && rows != 8) || interior_wgs < 0)
        return set_err(c, FWI_HIP_EINVAL, "set_pass_shape: need pass 0 or 1, rows 4 or 8, interior_wgs >= 0 (got %d, %d, %d)",
                       pass, rows, interior_wgs);
    int &r = pass ? c->strip_bwd : c->strip_fwd, &w = pass ? c->int_wgs_bwd : c->int_wgs_fwd;
    if (rows == r && interior_wgs == w) return FWI_HIP_OK;
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);  // captured launches hold the kernels, lists and LDS size
    w = interior_wgs;
    if (rows == r) return FWI_HIP_OK;
    r = rows;
    return c->nshots > 0 ? build_strip_lists(c) : FWI_HIP_OK;
}
extern "C" int fwi_hip_set_bwd_mixed(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_bwd_mixed: NULL context");
    drop_bwd(c);
    c->bwd_mixed = on != 0;
    return FWI_HIP_OK;
}
extern "C" int fwi_hip_set_fwd_mixed(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_fwd_mixed: NULL context");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);  // the checkpointed backward pass recomputes with the forward kernels
    c->fwd_mixed = on != 0;
    return FWI_HIP_OK;
}
extern "C" int fwi_hip_set_interior_wgs(fwi_hip_ctx *c, int wgs)
{
    if (!c || wgs < 0) return set_err(c, FWI_HIP_EINVAL, "set_interior_wgs: need wgs >= 0");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);  // captured launches hold the LDS size
    c->int_wgs_fwd = c->int_wgs_bwd = wgs;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_bwd_fused(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_bwd_fused: NULL context");
    drop_bwd(c);
    c->bwd_fused = on != 0;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_imaging(fwi_hip_ctx *c, int jump, int k0, int k1)
{
    if (!c || jump < 1 || k0 < 0 || k1 < k0) return set_err(c, FWI_HIP_EINVAL, "set_imaging: need jump >= 1, 0 <= k0 <= k1");
    if (jump != c->img_jump || k0 != c->img_k0 || k1 != c->img_k1)
        drop_bwd(c); // the z-marching kernels take their arguments by value: a captured graph holds the old window
    c->img_jump = jump;
    c->img_k0 = k0;
    c->img_k1 = k1;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_graph(fwi_hip_ctx *c, int steps)
{
    if (!c || steps < 0) return set_err(c, FWI_HIP_EINVAL, "set_graph: bad argument");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);
    c->graph_steps = steps;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_checkpointing(fwi_hip_ctx *c, int every, int storage)
{
    if (!c || every < 0 || storage < 0 || storage > 2)
        return set_err(c, FWI_HIP_EINVAL, "set_checkpointing: need every >= 0 and storage 0 (FP32), 1 (FP16) or 2 (BF16)");
    if (c->gx_fwd) (void)hipGraphExecDestroy(c->gx_fwd);
    c->gx_fwd = nullptr;
    drop_bwd(c);
    if (storage != c->ckpt_storage || every != c->ckpt_every) {
        dfree(c, c->ck_store);
        dfree(c, c->ck_scale);
        dfree(c, c->ck_cpml);
        dfree(c, c->ck_xseg);
        c->ck_store = nullptr;
        c->ck_scale = nullptr;
        c->ck_cpml = c->ck_xseg = nullptr;
        c->ck_nseg = c->ck_S = c->ck_K = 0;
    }
    c->ckpt_every = every;
    c->ckpt_storage = storage;
    c->rv_snaps = 0;
    c->have_fwd = false;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_revolve(fwi_hip_ctx *c, int snaps, int storage)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_revolve: NULL context");
    if (snaps < -1 || snaps > FWI_RV_MAX_SNAPS || storage < 0 || storage > 2)
        return set_err(c, FWI_HIP_EINVAL, "set_revolve: snaps must be -1..%d, storage 0 (FP32), 1 (FP16) or 2 (BF16)",
                       FWI_RV_MAX_SNAPS);
    if (storage != c->ckpt_storage) free_ckpt(c);
    c->rv_snaps = snaps;
    c->ckpt_storage = storage;
    if (snaps != 0) c->ckpt_every = 0;
    c->have_fwd = false;  // the next adjoint needs a forward pass with this plan
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_backward_plan(const fwi_hip_ctx *c, int *method, int *snaps, long long *recomputed)
{
    if (!c) return set_err(nullptr, FWI_HIP_EINVAL, "backward_plan: NULL context");
    if (method) *method = c->bwd_method;
    if (snaps) *snaps = c->bwd_snaps;
    if (recomputed) *recomputed = c->recomputed;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_plan_memory(const fwi_hip_desc *desc, int storage, int *method, int *snaps, size_t *bytes)
{
    if (!desc || !method || !snaps || storage < 0 || storage > 2)
        return set_err(nullptr, FWI_HIP_EINVAL, "plan_memory: bad argument");
    fwi_hip_ctx c;  // sizes only: no device call
    c.d = *desc;
    int rc = check_desc(&c, desc);
    if (rc) return rc;
    c.g = make_geom(desc);
    set_strides(&c);
    size_t base = 0;
    for (const auto &a : create_table(&c)) base += a.second;
    // set_shots: the source samples and one receiver row per shot
    base += sizeof(float) * (size_t)desc->max_shots * desc->max_nt + sizeof(int) * (size_t)desc->max_shots * c.g.ld;
    rc = plan_core(&c, base, storage, method, snaps);
    if (rc) {
        snprintf(g_err, sizeof(g_err), "%s", c.err);
        return rc;
    }
    if (bytes)
        *bytes = base + (*method == 0 ? frames_bytes(&c) : rv_fixed_bytes(&c) + (*snaps + 1) * rv_slot_bytes(&c, storage));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_revolve_schedule(int nt, int snaps, int *acts, int cap, int *nacts, long long *forward_steps)
{
    if (nt < 1 || snaps < 1 || cap < 0 || (cap > 0 && !acts) || !nacts)
        return set_err(nullptr, FWI_HIP_EINVAL, "revolve_schedule: bad argument");
    std::vector<RvAct> v;
    rv_schedule(nt, snaps, v);
    long long fsteps = 0;
    for (const RvAct &x : v) fsteps += x.op == RV_ADVANCE ? x.b - x.a : (x.op == RV_TURN ? 1 : 0);
    for (size_t i = 0; i < v.size() && i < (size_t)cap; i++) {
        acts[4 * i] = v[i].op;
        acts[4 * i + 1] = v[i].a;
        acts[4 * i + 2] = v[i].b;
        acts[4 * i + 3] = v[i].slot;
    }
    *nacts = (int)v.size();
    if (forward_steps) *forward_steps = fsteps;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_bwd_overlap(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_bwd_overlap: NULL context");
    drop_bwd(c);
    c->bwd_overlap = on != 0;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_bwd_chunk(fwi_hip_ctx *c, int shots)
{
    if (!c || shots < -1) return set_err(c, FWI_HIP_EINVAL, "set_bwd_chunk: need shots >= -1");
    drop_bwd(c);
    c->bwd_chunk = shots;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_bwd_general_rows(fwi_hip_ctx *c, int rows)
{
    if (!c || (rows != 4 && rows != 8)) return set_err(c, FWI_HIP_EINVAL, "set_bwd_general_rows: need 4 or 8");
    drop_bwd(c);
    c->gen_rows_bwd = rows;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_bwd_lanes(fwi_hip_ctx *c, int lanes)
{
    if (!c || lanes < 1 || lanes > 2) return set_err(c, FWI_HIP_EINVAL, "set_bwd_lanes: need 1 or 2");
    drop_bwd(c);
    c->bwd_lanes = lanes;
    return FWI_HIP_OK;
}

// tm: traces go to the host time-major (as the device records them) and fn,
// when given, sees each block of steps as soon as its rows have arrived
static int forward_impl(fwi_hip_ctx *c, float *traces, bool tm, fwi_hip_block_fn fn, void *user)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "forward: NULL context");
    if (!c->have_model || c->nshots == 0) return set_err(c, FWI_HIP_ESTATE, "forward: set_model and set_shots first");
    const FwiGeomDev &g = c->g;
    const long long S = c->nshots;
    // backward method: 1 uniform snapshots, 2 binomial checkpoints (fixed
    // slots, or from the HBM budget), else 0 boundary frames
    int method = c->ckpt_every > 0 ? 1 : 0, snaps = 0, rc = FWI_HIP_OK;
    if (method == 0 && c->rv_snaps > 0) {
        method = 2;
        snaps = c->rv_snaps;
    } else if (method == 0 && c->rv_snaps < 0) {
        if ((rc = plan_core(c, base_bytes(c), c->ckpt_storage, &method, &snaps))) return rc;
    }
    const bool ck = method != 0;
    if (ck && c->fwd_kind != 0) return set_err(c, FWI_HIP_EINVAL, "forward: checkpointing needs the tile kernels (set_kernels 0)");
    if (method == 0) {
        free_ckpt(c);
        rc = ensure_frames(c);
    } else {
        free_frames(c);
        rc = method == 1 ? ensure_ckpt(c) : ensure_slots(c, snaps + 1);
    }
    if (rc) return rc;
    c->bwd_method = method;
    c->bwd_snaps = snaps;
    if (method == 2 && (c->rv_sched_t != c->kmax || c->rv_sched_s != snaps)) {
        rv_schedule(c->kmax, snaps, c->rv_sched);
        c->rv_prefix = 0;
        while (c->rv_prefix < c->rv_sched.size() && c->rv_sched[c->rv_prefix].op != RV_TURN) c->rv_prefix++;
        c->rv_sched_t = c->kmax;
        c->rv_sched_s = snaps;
    }
    const int fmode = ck ? FWI_FWD_TRACES : FWI_FWD_FRAMES;
    bool pipe = false;  // trace blocks copied beside the steps
    HIPCHK(c, fwi_launch_fill_f32(c->P0, S * c->fs, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f32(c->P1, S * c->fs, 0.f, c->st));
    rc = zero_cpml(c);
    if (rc) return rc;
    FwiFwdArgs a;
    memset(&a, 0, sizeof(a));
    a.g = g;
    a.b = batch_dev(c);
    a.P0 = c->P0;
    a.P1 = c->P1;
    a.xh = c->xh;
    a.xd = c->xd;
    a.zh = c->zh;
    a.zd = c->zd;
    a.syn = c->syn_tm;
    a.pstrip = ck ? nullptr : c->pstrip;
    a.xstrip = ck ? nullptr : c->xstrip;
    c->fwd_args_h = a;  // host copy stays alive while the stream reads it
    HIPCHK(c, hipMemcpyAsync(c->fwd_args_d, &c->fwd_args_h, sizeof(FwiFwdArgs), hipMemcpyHostToDevice, c->st));
    const FwiFwdArgs *ad = c->fwd_args_d;
    const int nblocks = g.ntx * g.ntz * c->ngroups;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    const int B = c->graph_steps;
    hipEvent_t pe0;
    if ((rc = pass_begin(c, &pe0))) return rc;
    if (method == 1) {
        // a snapshot before every ckpt_every-th step, plain launches between
        const int K = c->ckpt_every;
        for (long long j = 0; j * K < c->kmax; j++) {
            if ((rc = ck_save(c, j))) return rc;
            for (int k = (int)(j * K); k < std::min((long long)c->kmax, (j + 1) * K); k++)
                HIPCHK(c, fwd_tile(c, &a, ad, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0), fmode));
        }
    } else if (method == 2) {
        // the binomial schedule's first sweep: S_0 in slot 0, its snapshots in
        // step order, and the state before the last step in slot `snaps`
        // (where the adjoint pass's first turn starts)
        size_t q = 0;
        for (int k = 0; k < c->kmax; k++) {
            if (k == 0 && (rc = ck_save_at(c, 0, 0, false))) return rc;
            for (; q < c->rv_prefix; q++) {
                const RvAct &x = c->rv_sched[q];
                if (x.op == RV_ADVANCE) continue;
                if (x.a != k) break;
                if ((rc = ck_save_at(c, x.slot, k, false))) return rc;
            }
            if (k == c->kmax - 1 && (rc = ck_save_at(c, snaps, k, false))) return rc;
            HIPCHK(c, fwd_tile(c, &a, ad, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0), fmode));
        }
    } else if (B > 0 && c->timing != 1) {
        if (!c->gx_fwd) {
            rc = capture(c, &c->gx_fwd, [&]() {
                hipError_t e = hipSuccess;
                for (int j = 0; j < B && e == hipSuccess; j++)
                    e = c->fwd_kind ? fwi_launch_fwd(&a, c->kbase, j, nblocks, c->st)
                                    : fwd_tile(c, &a, ad, c->tl_fwd, c->n_fwd, c->kbase, j, pair_of(c, 0), fmode);
                if (e == hipSuccess) e = fwi_launch_advance(c->kbase, B, c->st);
                return e;
            });
            if (rc) return rc;
        }
        HIPCHK(c, fwi_launch_set_int(c->kbase, 0, c->st));
        const int nblk = (c->kmax + B - 1) / B;
        pipe = tm && traces;
        if (pipe && (rc = ensure_copy(c, nblk))) return rc;
        // the copies go on st2 when the forward leaves it idle (one launch per
        // step, or the z-marching kernels): the copy stream may share a
        // hardware queue with st once the backward pass's four streams exist,
        // and copies queued behind the steps would no longer overlap them
        const bool st2_idle = c->fwd_kind != 0 || c->fwd_mixed || !c->concurrent;
        hipStream_t cs = st2_idle && c->st2 ? c->st2 : c->stc;
        for (int blk = 0; blk < nblk; blk++) {
            HIPCHK(c, hipGraphLaunch(c->gx_fwd, c->st));
            if (!pipe) continue;
            // the block's trace rows go to the host while the next block runs
            HIPCHK(c, hipEventRecord(c->blk_done[blk], c->st));
            HIPCHK(c, hipStreamWaitEvent(cs, c->blk_done[blk], 0));
            HIPCHK(c, copy_rows(c, traces, c->syn_tm, blk * B, std::min(c->kmax, (blk + 1) * B), true, cs));
            HIPCHK(c, hipEventRecord(c->blk_copy[blk], cs));
        }
    } else {
        for (int k = 0; k < c->kmax; k++) {
            hipEvent_t e0, e1;
            if ((rc = tick(c, &e0))) return rc;
            HIPCHK(c, c->fwd_kind ? fwi_launch_fwd(&a, nullptr, k, nblocks, c->st)
                                  : fwd_tile(c, &a, ad, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0), fmode));
            if ((rc = tick(c, &e1))) return rc;
            if (c->timing == 1) marks.push_back({0, {e0, e1}});
            if (marks.size() >= 512 && (rc = collect_timing(c, marks))) return rc;
        }
    }
    if ((rc = pass_end(c, 0, pe0, c->kmax))) return rc;
    if ((rc = collect_timing(c, marks))) return rc;
    for (int s = 0; s < c->nshots; s++) {
        const FwiShotDev &d = c->shots_h[s];
        if (d.nr == 0) continue;
        HIPCHK(c, fwi_launch_transpose(c->syn_tm + d.trace_off, c->syn + d.trace_off, d.nt, d.nr, c->st));
    }
    if (traces && !tm)
        HIPCHK(c, hipMemcpyAsync(traces, c->syn, sizeof(float) * c->ntrace, hipMemcpyDefault, c->st));
    if (traces && tm && !pipe)
        HIPCHK(c, hipMemcpyAsync(traces, c->syn_tm, sizeof(float) * c->ntrace, hipMemcpyDefault, c->st));
    int frc = 0;  // the first nonzero return of fn (later blocks are not handed over)
    if (pipe) {
        const int B = c->graph_steps, nblk = (c->kmax + B - 1) / B;
        for (int blk = 0; blk < nblk; blk++) {
            HIPCHK(c, hipEventSynchronize(c->blk_copy[blk]));
            if (fn && !frc) frc = fn(user, blk * B, std::min(c->kmax, (blk + 1) * B));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (tm && !pipe && fn) frc = fn(user, 0, c->kmax);
    c->have_fwd = true;
    c->have_res = false;
    c->res_host = nullptr;
    if (frc) return set_err(c, frc, "forward: the trace block callback returned %d", frc);
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_forward(fwi_hip_ctx *c, float *traces)
{
    return forward_impl(c, traces, false, nullptr, nullptr);
}

extern "C" int fwi_hip_forward_tm(fwi_hip_ctx *c, float *traces_tm, fwi_hip_block_fn fn, void *user)
{
    if (c && fn && !traces_tm) return set_err(c, FWI_HIP_EINVAL, "forward_tm: a callback needs the trace buffer");
    return forward_impl(c, traces_tm, true, fn, user);
}

extern "C" int fwi_hip_set_observed(fwi_hip_ctx *c, const float *obs)
{
    if (!c || !obs) return set_err(c, FWI_HIP_EINVAL, "set_observed: NULL argument");
    if (c->nshots == 0) return set_err(c, FWI_HIP_ESTATE, "set_observed: set_shots first");
    HIPCHK(c, hipMemcpyAsync(c->obs, obs, sizeof(float) * c->ntrace, hipMemcpyDefault, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_obs = true;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_misfit_l2(fwi_hip_ctx *c, double *misfit)
{
    if (!c || !misfit) return set_err(c, FWI_HIP_EINVAL, "misfit_l2: NULL argument");
    if (!c->have_fwd || !c->have_obs) return set_err(c, FWI_HIP_ESTATE, "misfit_l2: forward and set_observed first");
    const int nb = c->npartial;
    HIPCHK(c, fwi_launch_misfit_l2(c->syn, c->obs, c->res, c->ntrace, c->partial, nb, c->st));
    std::vector<double> part(nb);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->partial, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    double J = 0.0;
    for (double v : part) J += v;
    *misfit = J;
    c->have_res = true;
    c->res_host = nullptr;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_residual(fwi_hip_ctx *c, const float *res)
{
    if (!c || !res) return set_err(c, FWI_HIP_EINVAL, "set_residual: NULL argument");
    if (c->nshots == 0) return set_err(c, FWI_HIP_ESTATE, "set_residual: set_shots first");
    HIPCHK(c, hipMemcpyAsync(c->res, res, sizeof(float) * c->ntrace, hipMemcpyDefault, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_res = true;
    c->res_host = nullptr;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_residual_tm(fwi_hip_ctx *c, const float *res_tm)
{
    if (!c || !res_tm) return set_err(c, FWI_HIP_EINVAL, "set_residual_tm: NULL argument");
    if (c->nshots == 0) return set_err(c, FWI_HIP_ESTATE, "set_residual_tm: set_shots first");
    c->res_host = res_tm;  // read by the next adjoint, block by block beside its steps
    c->have_res = true;
    return FWI_HIP_OK;
}

// Shots per backward pass when the caller leaves it to the library
// (set_bwd_chunk -1, the default): a step of the backward pass reads the
// fields the step before wrote, so when the four wavefields (mu and p, two
// buffers each) of the chunks in flight fit in the device's last-level cache
// (256 MB Infinity Cache on MI355X; 200 MB of it counted) those reads come
// from it.  The chunks in flight must still give a step at least 10M
// cell-updates to fill the GPU.  At 1024x320 x 64 shots: 32 shots per pass
// with one chunk in flight (backward 148.1 -> 132.8 us per step), 16 with two
// (130.6 -> 119.6 us); 16 shots alone: 182 us.  At 4096x1024 x 16 shots no
// chunk fits.
#define FWI_LLC_USABLE 200.0e6
static int auto_bwd_chunk(const fwi_hip_ctx *c, int rgs, int lanes)
{
    const double shot_bytes = 4.0 * (double)c->fs * sizeof(float);
    const int n = (int)(FWI_LLC_USABLE / shot_bytes) / (rgs * lanes) * rgs;
    if (n <= 0 || n * lanes >= c->nshots) return 0;
    if ((double)n * lanes * c->g.pnx * c->g.pnz < 1.0e7) return 0;
    return n;
}

extern "C" int fwi_hip_adjoint(fwi_hip_ctx *c, double *grad, double *gk)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "adjoint: NULL context");
    if (!c->have_fwd || !c->have_res)
        return set_err(c, FWI_HIP_ESTATE, "adjoint: forward and a residual (misfit_l2 or set_residual) first");
    const FwiGeomDev &g = c->g;
    const long long S = c->nshots, nb = (long long)g.nx * g.nz;
    const bool tile = c->fwd_kind == 0;
    const int rgs = std::min(c->rec_group, c->nshots);        // tile path: shots per workgroup
    const int ngr = tile ? (c->nshots + rgs - 1) / rgs : c->ngroups;  // image accumulators in use
    HIPCHK(c, fwi_launch_fill_f32(c->M0, S * c->fs, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f32(c->M1, S * c->fs, 0.f, c->st));
    HIPCHK(c, fwi_launch_fill_f64(c->gacc, (long long)ngr * nb, 0.0, c->st));
    int rc = zero_cpml(c);
    if (rc) return rc;
    // a time-major residual from the host (set_residual_tm) goes up block by
    // block beside the steps in the graph form of a one-chunk boundary-saving
    // pass, else whole before them; a receiver-major one is transposed here
    const bool tm_res = c->res_host != nullptr;
    const bool pipe_res = tm_res && c->bwd_method == 0 && c->graph_steps > 0 && c->timing != 1;
    if (tm_res && !pipe_res)
        HIPCHK(c, hipMemcpyAsync(c->res_tm, c->res_host, sizeof(float) * c->ntrace, hipMemcpyDefault, c->st));
    for (int s = 0; s < c->nshots && !tm_res; s++) {
        const FwiShotDev &d = c->shots_h[s];
        if (d.nr == 0) continue;
        HIPCHK(c, fwi_launch_transpose(c->res + d.trace_off, c->res_tm + d.trace_off, d.nr, d.nt, c->st));
    }
    FwiAdjArgs a;
    memset(&a, 0, sizeof(a));
    a.g = g;
    a.b = batch_dev(c);
    a.M0 = c->M0;
    a.M1 = c->M1;
    a.xh = c->xh;
    a.xd = c->xd;
    a.zh = c->zh;
    a.zd = c->zd;
    a.res = c->res_tm;
    FwiRecArgs r;
    memset(&r, 0, sizeof(r));
    r.g = g;
    r.b = a.b;
    r.P0 = c->P0;
    r.P1 = c->P1;
    r.M0 = c->M0;
    r.M1 = c->M1;
    r.pstrip = c->pstrip;
    r.xstrip = c->xstrip;
    r.gacc = c->gacc;
    r.img_jump = c->img_jump;
    r.img_k0 = c->img_k0;
    r.img_k1 = c->img_k1;
    const bool fused_pass = tile && c->bwd_fused && c->ckpt_every == 0;
    if (fused_pass && c->bwd_mixed && !c->xg &&
        (rc = dalloc(c, (void **)&c->xg, sizeof(float) * 2 * (size_t)c->d.max_shots * nb)))
        return rc;
    r.xg = c->xg;
    r.xg_b = (long long)c->d.max_shots * nb;
    c->adj_args_h = a;
    c->rec_args_h = r;
    HIPCHK(c, hipMemcpyAsync(c->adj_args_d, &c->adj_args_h, sizeof(FwiAdjArgs), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->rec_args_d, &c->rec_args_h, sizeof(FwiRecArgs), hipMemcpyHostToDevice, c->st));
    const FwiAdjArgs *ad = c->adj_args_d;
    const FwiRecArgs *rd = c->rec_args_d;
    if (tile && c->rec_lists_for != rgs) {
        // items (group, tile) of the reconstruction kernel; interior when the tile is
        // interior for every shot of the group
        const int rr = c->strip_bwd;  // reconstruction rows per wave = the strip height
        const int ntz = (g.nz + rr - 1) / rr, ntiles = g.bntx * ntz;
        std::vector<int> lr[2];
        for (int q = 0; q < 2; q++) c->rec_pre[q].assign(ngr + 1, 0);
        for (int grp = 0; grp < ngr; grp++) {
            for (int q = 0; q < 2; q++) c->rec_pre[q][grp] = (int)lr[q].size();
            for (int t = 0; t < ntiles; t++) {
                const int bx0 = (t % g.bntx) * FWI_TX, z0 = (t / g.bntx) * rr;
                bool in = true;
                for (int s = grp * rgs; s < std::min((grp + 1) * rgs, c->nshots) && in; s++)
                    in = fwi_rec_tile_interior(g, c->shots_h[s], bx0, z0, rr);
                lr[in ? 0 : 1].push_back(grp * ntiles + t);
            }
        }
        for (int q = 0; q < 2; q++) c->rec_pre[q][ngr] = (int)lr[q].size();
        for (int q = 0; q < 2; q++) {
            c->n_rec[q] = (int)lr[q].size();
            if (c->n_rec[q])
                HIPCHK(c, hipMemcpyAsync(c->tl_rec[q], lr[q].data(), sizeof(int) * lr[q].size(), hipMemcpyHostToDevice,
                                         c->st));
        }
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->rec_lists_for = rgs;
    }
    const bool fused = tile && c->bwd_fused && c->ckpt_every == 0;
    const bool tall = tall_general(c);
    const int lists_key = rgs + (tall ? 1000000 : 0);
    if (fused && c->bwd_lists_for != lists_key) {
        // a tile of the padded grid is fused for a shot group when it is
        // interior for the adjoint and for the reconstruction of every shot of
        // the group; the other tiles run as adjoint strips + reconstruction
        // items.  The general side works on blocks of gr rows (the interior
        // tile height br, or 8 over 4-row interior tiles: a block is fused
        // when both of its 4-row tiles are, and its tiles then go to the
        // interior list; the fused test of 8 rows is that of its two halves)
        const int br = c->strip_bwd, gr = tall ? 8 : br;
        const int nsz = (g.pnz + br - 1) / br, ntiles = g.ntx * nsz;
        const int nszg = (g.pnz + gr - 1) / gr, ntg = g.ntx * nszg;
        std::vector<char> fz((size_t)ngr * ntg);
        std::vector<int> lb[4];
        for (int q = 0; q < 4; q++) c->bwd_pre[q].assign((q == 1 || q == 2 ? c->nshots : ngr) + 1, 0);
        for (int grp = 0; grp < ngr; grp++) {
            c->bwd_pre[0][grp] = (int)lb[0].size();
            c->bwd_pre[3][grp] = (int)lb[3].size();
            for (int t = 0; t < ntg; t++) {
                const int x0 = (t % g.ntx) * FWI_TX, r0 = (t / g.ntx) * gr;
                bool in = r0 + gr <= g.pnz;
                for (int s = grp * rgs; s < std::min((grp + 1) * rgs, c->nshots) && in; s++)
                    in = fwi_adj_strip_interior(g, c->shots_h[s], x0, r0, gr) &&
                         fwi_rec_tile_interior(g, c->shots_h[s], x0 - g.npml, r0, gr);
                fz[(size_t)grp * ntg + t] = in;
                if (!in && r0 < g.nz) lb[3].push_back(grp * ntg + t);
            }
            for (int t = 0; t < ntiles; t++) {   // interior tiles in row order
                const int tg = (t / g.ntx) * br / gr * g.ntx + t % g.ntx;
                if (fz[(size_t)grp * ntg + tg]) lb[0].push_back(grp * ntiles + t);
            }
        }
        c->bwd_pre[0][ngr] = (int)lb[0].size();
        c->bwd_pre[3][ngr] = (int)lb[3].size();
        // adjoint strips of the other blocks: [1] those whose rules are all on
        // the x side (interior strips included) or all on the z side with the
        // receiver rows, [2] the corner strips that meet both (the light and
        // the full body of k_adj_tile_gen)
        for (int s = 0; s < c->nshots; s++) {
            c->bwd_pre[1][s] = (int)lb[1].size();
            c->bwd_pre[2][s] = (int)lb[2].size();
            for (int t = 0; t < ntg; t++) {
                if (fz[(size_t)(s / rgs) * ntg + t]) continue;
                const int x0 = (t % g.ntx) * FWI_TX, sz = t / g.ntx;
                const int cls = fwi_adj_strip_class(g, c->shots_h[s], x0, sz * gr, gr);
                const bool corner = (cls & FWI_CLS_X) && (cls & (FWI_CLS_Z | FWI_CLS_RS));
                lb[corner ? 2 : 1].push_back((s * nszg + sz) * g.ntx + t % g.ntx);
            }
        }
        c->bwd_pre[1][c->nshots] = (int)lb[1].size();
        c->bwd_pre[2][c->nshots] = (int)lb[2].size();
        for (int q = 0; q < 4; q++) {
            c->n_bwd[q] = (int)lb[q].size();
            if (c->n_bwd[q])
                HIPCHK(c, hipMemcpyAsync(c->tl_bwd[q], lb[q].data(), sizeof(int) * lb[q].size(), hipMemcpyHostToDevice,
                                         c->st));
        }
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->bwd_lists_for = lists_key;
        c->bwd_work_for = -1;
    }
    const int nb_adj = g.ntx * g.ntz * c->ngroups;
    const int nb_rec = g.bntx * g.bntz * c->ngroups;
    hipEvent_t pe0;
    if ((rc = pass_begin(c, &pe0))) return rc;
    if (c->bwd_method != 0) {
        if (!tile) return set_err(c, FWI_HIP_EINVAL, "adjoint: checkpointing needs the tile kernels (set_kernels 0)");
        if (!c->ck_store) return set_err(c, FWI_HIP_ESTATE, "adjoint: no snapshots (forward with checkpointing first)");
        // per segment from the last: restore its snapshot, recompute its
        // steps storing X, then run its adjoint steps imaging mu X
        const long long nbb = (long long)g.nx * g.nz, K = c->ckpt_every;
        FwiFwdArgs f;
        memset(&f, 0, sizeof(f));
        f.g = g;
        f.b = a.b;
        f.P0 = c->P0;
        f.P1 = c->P1;
        f.xh = c->xh2;
        f.xd = c->xd2;
        f.zh = c->zh2;
        f.zd = c->zd2;
        f.xseg = c->ck_xseg;
        f.xseg_step = S * nbb;
        c->fwd_rc_h = f;
        HIPCHK(c, hipMemcpyAsync(c->fwd_rc_d, &c->fwd_rc_h, sizeof(FwiFwdArgs), hipMemcpyHostToDevice, c->st));
        int *k0_dev = &c->fwd_rc_d->xseg_k0;  // device address of the segment's first step
        long long rec = 0;
        if (c->bwd_method == 2) {
            // the rest of the binomial schedule, from the state before the last step
            if ((rc = ck_restore_at(c, c->bwd_snaps, c->kmax - 1))) return rc;
            for (size_t i = c->rv_prefix; i < c->rv_sched.size(); i++) {
                const RvAct &x = c->rv_sched[i];
                if (x.op == RV_ADVANCE) {
                    for (int k = x.a; k < x.b; k++)
                        HIPCHK(c, fwd_tile(c, &f, c->fwd_rc_d, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0),
                                                      FWI_FWD_PLAIN));
                    rec += x.b - x.a;
                } else if (x.op == RV_TAKESHOT) {
                    if ((rc = ck_save_at(c, x.slot, x.a, true))) return rc;
                } else if (x.op == RV_RESTORE) {
                    if ((rc = ck_restore_at(c, x.slot, x.a))) return rc;
                } else {  // turn: X^k from S_k, then the adjoint step k and its image
                    const int k = x.a;
                    HIPCHK(c, fwi_launch_set_int(k0_dev, k, c->st));
                    HIPCHK(c, fwd_tile(c, &f, c->fwd_rc_d, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0),
                                                  FWI_FWD_XSEG));
                    HIPCHK(c, adj_tile(c, &a, ad, c->tl_adj, c->n_adj, nullptr, k, pair_of(c, 1)));
                    HIPCHK(c, fwi_launch_img_seg(&r, rd, c->ck_xseg, k, rgs, ngr, c->st));
                    rec++;
                }
            }
        }
        // imaging step k (mu^{k+1} X^k, mu^{k+1} in M[(k+1) & 1]) on a stream of
        // its own beside the adjoint step k - 1, which only reads that buffer;
        // the adjoint step k - 2 overwrites it, so it waits for the imaging
        // step k, and a segment's recompute, which overwrites the stored X,
        // waits for the previous segment's last imaging step.  The images of
        // a group still add up step by step in the same order.
        hipStream_t si = c->concurrent && c->stB ? c->stB : c->st;
        int last_img = -1;  // the last imaging step queued on si
        for (long long j = c->bwd_method == 1 ? (c->kmax + K - 1) / K - 1 : -1; j >= 0; j--) {
            const int k0 = (int)(j * K), k1 = (int)std::min((long long)c->kmax, (j + 1) * K);
            if (si != c->st && last_img >= 0) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_img[last_img & 1], 0));
            if ((rc = ck_restore(c, j))) return rc;
            rec += k1 - k0;
            HIPCHK(c, fwi_launch_set_int(k0_dev, k0, c->st));
            for (int k = k0; k < k1; k++)
                HIPCHK(c, fwd_tile(c, &f, c->fwd_rc_d, c->tl_fwd, c->n_fwd, nullptr, k, pair_of(c, 0),
                                              FWI_FWD_XSEG));
            for (int k = k1 - 1; k >= k0; k--) {
                // imaging step k + 2 (queued when k + 2 < kmax) reads the buffer
                // this adjoint step writes; ev_img[k & 1] was last recorded by it
                if (si != c->st && k + 2 < c->kmax) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_img[k & 1], 0));
                HIPCHK(c, adj_tile(c, &a, ad, c->tl_adj, c->n_adj, nullptr, k, pair_of(c, 1)));
                if (si != c->st) {
                    HIPCHK(c, hipEventRecord(c->ev_adjd, c->st));
                    HIPCHK(c, hipStreamWaitEvent(si, c->ev_adjd, 0));
                }
                HIPCHK(c, fwi_launch_img_seg(&r, rd, c->ck_xseg + (long long)(k - k0) * S * nbb, k, rgs, ngr, si));
                if (si != c->st) HIPCHK(c, hipEventRecord(c->ev_img[k & 1], si));
                last_img = k;
            }
        }
        if (si != c->st && last_img >= 0) HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_img[last_img & 1], 0));
        c->recomputed = rec;
        if ((rc = pass_end(c, 1, pe0, c->kmax))) return rc;
        HIPCHK(c, fwi_launch_grad_finalize(c->gacc, ngr, nb, c->vel, c->gk, c->grad, c->st));
        if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->grad, sizeof(double) * nb, hipMemcpyDefault, c->st));
        if (gk) HIPCHK(c, hipMemcpyAsync(gk, c->gk, sizeof(double) * nb, hipMemcpyDefault, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->have_fwd = false;
        c->res_host = nullptr;
        return FWI_HIP_OK;
    }
    // shot chunks, each a whole backward pass over its own slice of the
    // strip and item lists (shot-major and group-major), so that a chunk's
    // working set (mu, p and the FP64 images of its shots) can stay cached
    // Two chunks in flight (bwd_lanes 2, graph form of the fused step): chunk
    // ch + 1 runs its steps on stB/stB2 beside chunk ch on st/st2, so that each
    // fills the other's joins and tails; the two touch disjoint shots.
    const bool lanes_ok = fused && !c->bwd_mixed && c->graph_steps > 0 && c->timing != 1 && c->concurrent;
    const int lanes = lanes_ok ? c->bwd_lanes : 1;
    const int chunk = c->bwd_chunk < 0 ? auto_bwd_chunk(c, rgs, lanes) : c->bwd_chunk;
    const int cs = (tile && chunk > 0 && chunk < c->nshots) ? ((chunk + rgs - 1) / rgs) * rgs : c->nshots;
    const int nch = (c->nshots + cs - 1) / cs;
    if ((int)c->gx_bwd.size() != nch) {
        drop_bwd(c);
        c->gx_bwd.assign(nch, nullptr);
    }
    // the one-launch backward step: the workgroup order of every chunk, the
    // general roles first (reconstruction, corners, x/z-side strips)
    const bool mixed = fused && c->bwd_mixed && c->xg;
    const long long work_key = (long long)rgs * 1000003LL + cs;
    if (mixed && c->bwd_work_for != work_key) {
        std::vector<int> all;
        c->bwd_work_off.assign(nch, 0);
        c->bwd_work_n.assign(nch, 0);
        for (int ch = 0; ch < nch; ch++) {
            const int s0 = ch * cs, s1 = std::min(s0 + cs, c->nshots), g0 = s0 / rgs, g1 = (s1 + rgs - 1) / rgs;
            int n[4];
            for (int q = 0; q < 4; q++) {
                const bool per_shot = q == 1 || q == 2;
                n[q] = c->bwd_pre[q][per_shot ? s1 : g1] - c->bwd_pre[q][per_shot ? s0 : g0];
            }
            const std::vector<int> w = work_order({n[0], (n[1] + 3) / 4, (n[2] + 3) / 4, n[3]}, {3, 2, 1, 0});
            c->bwd_work_off[ch] = (long long)all.size();
            c->bwd_work_n[ch] = (int)w.size();
            all.insert(all.end(), w.begin(), w.end());
        }
        if ((long long)all.size() > c->bwd_work_cap) {
            dfree(c, c->bwd_work);
            c->bwd_work = nullptr;
            c->bwd_work_cap = 0;
            if ((rc = dalloc(c, (void **)&c->bwd_work, sizeof(int) * all.size()))) return rc;
            c->bwd_work_cap = (long long)all.size();
        }
        if (!all.empty())
            HIPCHK(c, hipMemcpyAsync(c->bwd_work, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->bwd_work_for = work_key;
    }
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    const int B = c->graph_steps;
    for (int ch = 0; ch < nch;) {
        const bool pair = lanes == 2 && ch + 1 < nch;  // chunks ch and ch + 1 in flight
        const int s0 = ch * cs, s1 = std::min(s0 + cs, c->nshots), s2 = pair ? std::min(s1 + cs, c->nshots) : s1;
        int kmax = 0;
        for (int s = s0; s < s2; s++) kmax = std::max(kmax, c->shots_h[s].nt);
        const int *ta[2], *tr[2], *tb[4] = {nullptr, nullptr, nullptr, nullptr};
        int na[2], nr[2], nbw[4] = {0, 0, 0, 0};
        if (tile) {
            const int g0 = s0 / rgs, g1 = (s1 + rgs - 1) / rgs;
            for (int q = 0; q < 2; q++) {
                ta[q] = c->tl_adj[q] + c->adj_pre[q][s0];
                na[q] = c->adj_pre[q][s1] - c->adj_pre[q][s0];
                tr[q] = c->tl_rec[q] + c->rec_pre[q][g0];
                nr[q] = c->rec_pre[q][g1] - c->rec_pre[q][g0];
            }
            if (fused) {
                for (int q = 0; q < 4; q++) {
                    const bool per_shot = q == 1 || q == 2;
                    const int lo = per_shot ? s0 : g0, hi = per_shot ? s1 : g1;
                    tb[q] = c->tl_bwd[q] + c->bwd_pre[q][lo];
                    nbw[q] = c->bwd_pre[q][hi] - c->bwd_pre[q][lo];
                }
            }
        }
        // the second chunk's lists (shots s1 .. s2 - 1)
        const int *tb2[4] = {nullptr, nullptr, nullptr, nullptr};
        int nbw2[4] = {0, 0, 0, 0};
        if (pair) {
            const int g1 = s1 / rgs, g2 = (s2 + rgs - 1) / rgs;
            for (int q = 0; q < 4; q++) {
                const bool per_shot = q == 1 || q == 2;
                const int lo = per_shot ? s1 : g1, hi = per_shot ? s2 : g2;
                tb2[q] = c->tl_bwd[q] + c->bwd_pre[q][lo];
                nbw2[q] = c->bwd_pre[q][hi] - c->bwd_pre[q][lo];
            }
        }
        auto launch_adj = [&](const int *kb, int koff) {
            return tile ? adj_tile(c, &a, ad, ta, na, kb, koff, pair_of(c, 1))
                        : fwi_launch_adj(&a, kb, koff, nb_adj, c->st);
        };
        auto launch_rec = [&](const int *kb, int koff) {
            return tile ? rec_tile(c, &r, rd, tr, nr, kb, koff, rgs, pair_of(c, 1))
                        : fwi_launch_rec(&r, kb, koff, nb_rec, c->st);
        };
        auto launch_bwd = [&](const int *kb, int koff) {
            if (mixed) {
                const int *w = c->bwd_work + c->bwd_work_off[ch];
                return c->strip_bwd == 4
                           ? fwi_launch_bwd_mixed_s4(&a, ad, rd, w, c->bwd_work_n[ch], tb, nbw, kb, koff, rgs, c->st)
                           : fwi_launch_bwd_mixed(&a, ad, rd, w, c->bwd_work_n[ch], tb, nbw, kb, koff, rgs, c->st);
            }
            return bwd_tile(c, &a, ad, rd, tb, nbw, kb, koff, rgs, pair_of(c, 1));
        };
        // overlapped form: adjoint k on st, imaging k+1 on st2, joined per step;
        // k runs down to -1 so that the last pair images step 0 (kernels
        // return for k < 0 and for k >= a shot's nt)
        const bool ovl = tile && c->bwd_overlap && c->timing != 1 && !fused;
        auto pair_step = [&](const int *kb, int koff) {
            FwiPair p1 = {c->st, nullptr, c->ev_fork, c->ev_join, 0},
                    p2 = {c->st2, nullptr, c->ev_fork, c->ev_join, 0};
            hipError_t e = hipEventRecord(c->ev_fork2, c->st);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->st2, c->ev_fork2, 0);
            if (e == hipSuccess) e = rec_tile(c, &r, rd, tr, nr, kb, koff + 1, rgs, p2);
            if (e == hipSuccess) e = adj_tile(c, &a, ad, ta, na, kb, koff, p1);
            if (e == hipSuccess) e = hipEventRecord(c->ev_join2, c->st2);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->st, c->ev_join2, 0);
            return e;
        };
        if (B > 0 && c->timing != 1) {
            if (!c->gx_bwd[ch]) {
                rc = capture(c, &c->gx_bwd[ch], [&]() {
                    hipError_t e = hipSuccess;
                    if (pair) {
                        FwiPair p2 = pair_of(c, 1);
                        p2.st = c->stB;
                        p2.st2 = c->stB2;
                        p2.fork = c->evB_fork;
                        p2.join = c->evB_join;
                        e = hipEventRecord(c->evL_fork, c->st);
                        if (e == hipSuccess) e = hipStreamWaitEvent(c->stB, c->evL_fork, 0);
                        for (int j = 0; j < B && e == hipSuccess; j++) {
                            e = launch_bwd(c->kbase, -j);
                            if (e == hipSuccess) e = bwd_tile(c, &a, ad, rd, tb2, nbw2, c->kbase, -j, rgs, p2);
                        }
                        if (e == hipSuccess) e = hipEventRecord(c->evL_join, c->stB);
                        if (e == hipSuccess) e = hipStreamWaitEvent(c->st, c->evL_join, 0);
                        if (e == hipSuccess) e = fwi_launch_advance(c->kbase, -B, c->st);
                        return e;
                    }
                    for (int j = 0; j < B && e == hipSuccess; j++) {
                        if (fused) {
                            e = launch_bwd(c->kbase, -j);
                        } else if (ovl) {
                            e = pair_step(c->kbase, -j);
                        } else {
                            e = launch_adj(c->kbase, -j);
                            if (e == hipSuccess) e = launch_rec(c->kbase, -j);
                        }
                    }
                    if (e == hipSuccess) e = fwi_launch_advance(c->kbase, -B, c->st);
                    return e;
                });
                if (rc) return rc;
            }
            HIPCHK(c, fwi_launch_set_int(c->kbase, kmax - 1, c->st));
            const int nblk = ovl ? (kmax + B) / B : (kmax + B - 1) / B;
            // the residual goes up beside the first chunk's blocks (every
            // shot's rows, so the later chunks find theirs in place); rows of
            // steps past the first chunk's last one first
            const bool up = pipe_res && ch == 0;
            if (up && (rc = ensure_copy(c, nblk + 1))) return rc;
            if (up && kmax < c->kmax) {
                HIPCHK(c, copy_rows(c, const_cast<float *>(c->res_host), c->res_tm, kmax, c->kmax, false, c->stc));
                HIPCHK(c, hipEventRecord(c->blk_copy[nblk], c->stc));
                HIPCHK(c, hipStreamWaitEvent(c->st, c->blk_copy[nblk], 0));
            }
            for (int blk = 0; blk < nblk; blk++) {
                if (up) {
                    // residual rows of the block's steps kmax-1-blk*B down to kmax-(blk+1)*B
                    const int hi = kmax - blk * B, lo = std::max(0, hi - B);
                    if (hi > 0)
                        HIPCHK(c, copy_rows(c, const_cast<float *>(c->res_host), c->res_tm, lo, hi, false, c->stc));
                    HIPCHK(c, hipEventRecord(c->blk_copy[blk], c->stc));
                    HIPCHK(c, hipStreamWaitEvent(c->st, c->blk_copy[blk], 0));
                }
                HIPCHK(c, hipGraphLaunch(c->gx_bwd[ch], c->st));
            }
            // the one-launch step images its general tiles one step late:
            // the image of step 0 follows the last step (the last block ran
            // k = -1 itself when it went past step 0)
            if (mixed && kmax - nblk * B >= 0) HIPCHK(c, launch_bwd(nullptr, -1));
        } else if (ovl) {
            for (int k = kmax - 1; k >= -1; k--) HIPCHK(c, pair_step(nullptr, k));
        } else if (fused) {
            for (int k = kmax - 1; k >= 0; k--) {
                hipEvent_t e0, e1;
                if ((rc = tick(c, &e0))) return rc;
                HIPCHK(c, launch_bwd(nullptr, k));
                if ((rc = tick(c, &e1))) return rc;
                if (c->timing == 1) marks.push_back({1, {e0, e1}});
                if (marks.size() >= 512 && (rc = collect_timing(c, marks))) return rc;
            }
            if (mixed) HIPCHK(c, launch_bwd(nullptr, -1));  // the image of step 0
        } else {
            for (int k = kmax - 1; k >= 0; k--) {
                hipEvent_t e0, e1, e2;
                if ((rc = tick(c, &e0))) return rc;
                HIPCHK(c, launch_adj(nullptr, k));
                if ((rc = tick(c, &e1))) return rc;
                HIPCHK(c, launch_rec(nullptr, k));
                if ((rc = tick(c, &e2))) return rc;
                if (c->timing == 1) {
                    marks.push_back({1, {e0, e1}});
                    marks.push_back({2, {e1, e2}});
                }
                if (marks.size() >= 512 && (rc = collect_timing(c, marks))) return rc;
            }
        }
        ch += pair ? 2 : 1;
    }
    if ((rc = pass_end(c, 1, pe0, c->kmax))) return rc;
    if ((rc = collect_timing(c, marks))) return rc;
    HIPCHK(c, fwi_launch_grad_finalize(c->gacc, ngr, nb, c->vel, c->gk, c->grad, c->st));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->grad, sizeof(double) * nb, hipMemcpyDefault, c->st));
    if (gk) HIPCHK(c, hipMemcpyAsync(gk, c->gk, sizeof(double) * nb, hipMemcpyDefault, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    // the reconstruction ran the forward's final state back to the start:
    // another adjoint needs another forward
    c->have_fwd = false;
    c->res_host = nullptr;
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_synchronize(fwi_hip_ctx *c)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "synchronize: NULL context");
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_set_timing(fwi_hip_ctx *c, int on)
{
    if (!c) return set_err(c, FWI_HIP_EINVAL, "set_timing: NULL context");
    if (on < 0 || on > 2) return set_err(c, FWI_HIP_EINVAL, "set_timing: mode 0, 1 or 2");
    c->timing = on;
    for (int i = 0; i < 3; i++) {
        c->t_ms[i] = 0;
        c->t_n[i] = 0;
    }
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_kernel_time(fwi_hip_ctx *c, int which, double *ms, long long *launches)
{
    if (!c || which < 0 || which > 2) return set_err(c, FWI_HIP_EINVAL, "kernel_time: bad argument");
    if (ms) *ms = c->t_ms[which];
    if (launches) *launches = c->t_n[which];
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_read_state(fwi_hip_ctx *c, int what, int shot, float *out)
{
    if (!c || !out || shot < 0 || shot >= c->nshots || what < 0 || what > 5)
        return set_err(c, FWI_HIP_EINVAL, "read_state: bad argument");
    const FwiGeomDev &g = c->g;
    if (what <= 3) {
        float *base[] = {c->P0, c->P1, c->M0, c->M1};
        const float *src = base[what] + (long long)shot * c->fs;
        HIPCHK(c, hipMemcpy2DAsync(out, sizeof(float) * g.pnx, src, sizeof(float) * g.ld, sizeof(float) * g.pnx,
                                   g.pnz, hipMemcpyDeviceToHost, c->st));
    } else {
        const FwiShotDev &d = c->shots_h[shot];
        if (!c->pstrip) return set_err(c, FWI_HIP_ESTATE, "read_state: no boundary frames (checkpointed forward)");
        const float *src = (what == 4 ? c->pstrip : c->xstrip) + d.strip_off;
        HIPCHK(c, hipMemcpyAsync(out, src, sizeof(float) * (size_t)d.nt * g.sl, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return FWI_HIP_OK;
}

// --------------------------------------------------------------------------
// per-step entry points (caller-owned device buffers)

static int ref_args(const fwi_hip_grid *g, FwiRefArgs &a, int pnx, int pnz)
{
    if (!g) return set_err(nullptr, FWI_HIP_EINVAL, "NULL grid");
    int rc = check_grid(nullptr, *g);
    if (rc) return rc;
    memset(&a, 0, sizeof(a));
    Coe co = fd_coe(g->dx, g->dz);
    a.pnx = pnx;
    a.pnz = pnz;
    a.npml = g->npml;
    a.c0x = co.c0x; a.c1x = co.c1x; a.ix = co.ix;
    a.c0z = co.c0z; a.c1z = co.c1z; a.iz = co.iz;
    a.iz840 = co.iz840; a.iz24 = co.iz24;
    return FWI_HIP_OK;
}

#define HIPCHK0(call)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_err(nullptr, FWI_HIP_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
    } while (0)

extern "C" int fwi_hip_op_pml(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux, float *vz,
                              float *u0, float *u1, float *v1, float *ux0, float *ux1, float *vz1, void *stream)
{
    FwiRefArgs a;
    int rc = ref_args(g, a, g ? g->nx + 2 * g->npml : 0, g ? g->nz + g->npml : 0);
    if (rc) return rc;
    if (!p || !u || !v || !ux || !vz || !u0 || !u1 || !v1 || !ux0 || !ux1 || !vz1)
        return set_err(nullptr, FWI_HIP_EINVAL, "op_pml: NULL buffer");
    hipStream_t st = (hipStream_t)stream;
    std::vector<float> tab;
    cpml_tables(*g, tab);
    float *dtab = nullptr;
    HIPCHK0(hipMalloc(&dtab, sizeof(float) * tab.size()));
    hipError_t e = hipMemcpy(dtab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice);
    a.p = p; a.u = u; a.v = v; a.ux = ux; a.vz = vz;
    a.u0 = u0; a.u1 = u1; a.v1 = v1; a.ux0 = ux0; a.ux1 = ux1; a.vz1 = vz1;
    a.tab = dtab;
    for (int ph = 0; ph < 12 && e == hipSuccess; ph++) e = fwi_launch_ref_phase(&a, ph, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dtab);
    if (e != hipSuccess) return set_err(nullptr, FWI_HIP_EHIP, "op_pml: %s", hipGetErrorString(e));
    return FWI_HIP_OK;
}

// For the allocation-free entry point (fwi_noalloc.hip): the argument block of
// the reference-structured a_lap_opt_2d_pml phases, the CPML tables on the
// host (tab[6 (2 npml + 1)]) and the library's error message.
extern "C" int fwi_hip_internal_op_pml_args(const fwi_hip_grid *g, FwiRefArgs *a, float *tab)
{
    if (!a || !tab) return set_err(nullptr, FWI_HIP_EINVAL, "op_pml: NULL argument block");
    int rc = ref_args(g, *a, g ? g->nx + 2 * g->npml : 0, g ? g->nz + g->npml : 0);
    if (rc) return rc;
    std::vector<float> t;
    cpml_tables(*g, t);
    memcpy(tab, t.data(), sizeof(float) * t.size());
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_internal_err(int code, const char *msg)
{
    return set_err(nullptr, code, "%s", msg);
}

extern "C" int fwi_hip_op_interior(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux, float *vz,
                                   void *stream)
{
    FwiRefArgs a;
    int rc = ref_args(g, a, g ? g->nx : 0, g ? g->nz : 0);
    if (rc) return rc;
    if (!p || !u || !v || !ux || !vz) return set_err(nullptr, FWI_HIP_EINVAL, "op_interior: NULL buffer");
    a.p = p; a.u = u; a.v = v; a.ux = ux; a.vz = vz;
    for (int ph = 20; ph < 24; ph++) HIPCHK0(fwi_launch_ref_phase(&a, ph, (hipStream_t)stream));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_time_integral(float *pnew, const float *pnow, const float *pold, const float *pxx,
                                     const float *pzz, const float *ekp, size_t m, size_t n, void *stream)
{
    if (!pnew || !pnow || !pold || !pxx || !pzz || !ekp) return set_err(nullptr, FWI_HIP_EINVAL, "time_integral: NULL");
    HIPCHK0(fwi_launch_ref_time_integral(pnew, pnow, pold, pxx, pzz, ekp, (long long)(m * n), (hipStream_t)stream));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_write_bdr(const fwi_hip_grid *g, float *ud, float *lr, const float *pf, void *stream)
{
    if (!g || !ud || !lr || !pf) return set_err(nullptr, FWI_HIP_EINVAL, "write_bdr: NULL");
    HIPCHK0(fwi_launch_ref_write_bdr(ud, lr, pf, g->nx, g->nz, g->npml, 0, (hipStream_t)stream));
    return FWI_HIP_OK;
}

// write_bdr_fs_no (fwi_2d.c:962-997): the CPML also above the physical grid
extern "C" int fwi_hip_write_bdr_fs_no(const fwi_hip_grid *g, float *ud, float *lr, const float *pf, void *stream)
{
    if (!g || !ud || !lr || !pf) return set_err(nullptr, FWI_HIP_EINVAL, "write_bdr_fs_no: NULL");
    HIPCHK0(fwi_launch_ref_write_bdr(ud, lr, pf, g->nx, g->nz, g->npml, g->npml, (hipStream_t)stream));
    return FWI_HIP_OK;
}

// model_ext_fs_no (fwi_2d.c:29-65, host): model[n2][n1] into the middle of
// emodel[en2][en1], the margins (en - n)/2 wide on every side filled with
// the nearest edge or corner value
extern "C" int fwi_hip_model_ext_fs_no(const float *model, int n1, int n2, float *emodel, int en1, int en2)
{
    if (!model || !emodel || n1 < 1 || n2 < 1 || en1 < n1 || en2 < n2)
        return set_err(nullptr, FWI_HIP_EINVAL, "model_ext_fs_no: NULL argument or extended grid smaller");
    if ((en1 - n1) % 2 || (en2 - n2) % 2)
        return set_err(nullptr, FWI_HIP_EINVAL,
                       "model_ext_fs_no: margins must be equal on both sides (en1 - n1 = %d, en2 - n2 = %d)", en1 - n1,
                       en2 - n2);
    const int g1 = (en1 - n1) / 2, g2 = (en2 - n2) / 2;
    for (int i2 = 0; i2 < en2; i2++) {
        const int s2 = std::min(std::max(i2 - g2, 0), n2 - 1);
        for (int i1 = 0; i1 < en1; i1++) {
            const int s1 = std::min(std::max(i1 - g1, 0), n1 - 1);
            emodel[(size_t)i2 * en1 + i1] = model[(size_t)s2 * n1 + s1];
        }
    }
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_read_bdr(const fwi_hip_grid *g, const float *ud, const float *lr, float *pb, void *stream)
{
    if (!g || !ud || !lr || !pb) return set_err(nullptr, FWI_HIP_EINVAL, "read_bdr: NULL");
    HIPCHK0(fwi_launch_ref_read_bdr(ud, lr, pb, g->nx, g->nz, (hipStream_t)stream));
    return FWI_HIP_OK;
}

// --------------------------------------------------------------------------
// device memory helpers

extern "C" int fwi_hip_device_count(int *n)
{
    if (!n) return set_err(nullptr, FWI_HIP_EINVAL, "device_count: NULL");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        (void)hipGetLastError();
    }
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_dmalloc(void **ptr, size_t bytes)
{
    if (!ptr) return set_err(nullptr, FWI_HIP_EINVAL, "dmalloc: NULL");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 4);
    if (e != hipSuccess) return set_err(nullptr, FWI_HIP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemset(*ptr, 0, bytes ? bytes : 4);
    if (e != hipSuccess) return set_err(nullptr, FWI_HIP_EHIP, "hipMemset: %s", hipGetErrorString(e));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_dfree(void *ptr)
{
    if (ptr) HIPCHK0(hipFree(ptr));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_host_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return set_err(nullptr, FWI_HIP_EINVAL, "host_alloc: NULL");
    *ptr = nullptr;
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 4, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return set_err(nullptr, FWI_HIP_ENOMEM, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_host_free(void *ptr)
{
    if (ptr) HIPCHK0(hipHostFree(ptr));
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    if (!dst || !src || kind < 1 || kind > 3) return set_err(nullptr, FWI_HIP_EINVAL, "memcpy: bad argument");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIPCHK0(hipDeviceSynchronize());
    HIPCHK0(hipMemcpy(dst, src, bytes, k));
    return FWI_HIP_OK;
}
