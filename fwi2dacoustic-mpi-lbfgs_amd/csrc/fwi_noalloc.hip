// fwi_noalloc.hip -- fwi_hip_op_pml_noalloc: the reference-structured
// a_lap_opt_2d_pml step (fwi_2d.c:761-864) on caller-owned device buffers,
// allocating no device memory, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// fwi_hip_op_pml allocates a device buffer for the CPML coefficient tables on
// every call.  Here the tables go to `scratch`, a static device array of the
// library: one copy per GPU, held until the process ends.  Its address on the
// current device comes from hipGetSymbolAddress, once per device, and only
// that address is handed to kernels (in host code the array's name is a
// host-side stand-in, not a device address).  The tables reach the array
// through kernel arguments, so no host buffer has to outlive the call and no
// copy node refers to pageable memory.
//
// Every call on a GPU rewrites the same array: calls that overlap on one GPU
// (different streams or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fwi_dev.h"
#include "../../include/fwi_hip.h"

extern "C" {
int fwi_hip_internal_op_pml_args(const fwi_hip_grid *g, FwiRefArgs *a, float *tab);
int fwi_hip_internal_err(int code, const char *msg);
hipError_t fwi_launch_ref_phase(const FwiRefArgs *a, int phase, hipStream_t st);
}

// the largest CPML the library accepts (npml <= 64): 6 rows of 2 npml + 1 floats
#define FWI_NOALLOC_MAXNPML 64
#define FWI_NOALLOC_TABN (6 * (2 * FWI_NOALLOC_MAXNPML + 1))
static __device__ __align__(256) float scratch[FWI_NOALLOC_TABN];

// tables travel in kernel arguments, FWI_NOALLOC_CHUNK floats (1 KiB) per launch
#define FWI_NOALLOC_CHUNK 256
struct TabChunk {
    float v[FWI_NOALLOC_CHUNK];
};

__global__ void k_tab_chunk(TabChunk c, int n, float *dst)
{
    const int i = threadIdx.x;
    if (i < n) dst[i] = c.v[i];
}

#define FWI_NOALLOC_MAXDEV 64
static std::mutex g_mu;
static float *g_scratch[FWI_NOALLOC_MAXDEV];

// address of `scratch` on the current device (looked up once per device)
static int scratch_ptr(float **out)
{
    *out = nullptr;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= FWI_NOALLOC_MAXDEV)
        return fwi_hip_internal_err(FWI_HIP_EHIP, "op_pml_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fwi_hip_internal_err(FWI_HIP_EHIP, "op_pml_noalloc: hipGetSymbolAddress(scratch) failed");
        g_scratch[dev] = (float *)p;
    }
    *out = g_scratch[dev];
    return FWI_HIP_OK;
}

extern "C" int fwi_hip_op_pml_noalloc(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux,
                                      float *vz, float *u0, float *u1, float *v1, float *ux0, float *ux1,
                                      float *vz1, void *stream)
{
    if (!g) return fwi_hip_internal_err(FWI_HIP_EINVAL, "op_pml_noalloc: NULL grid");
    if (g->npml > FWI_NOALLOC_MAXNPML)
        return fwi_hip_internal_err(FWI_HIP_EINVAL,
                                    "op_pml_noalloc: npml exceeds 64 (the CPML tables' static device array "
                                    "holds 6 x 129 floats)");
    if (!p || !u || !v || !ux || !vz || !u0 || !u1 || !v1 || !ux0 || !ux1 || !vz1)
        return fwi_hip_internal_err(FWI_HIP_EINVAL, "op_pml_noalloc: NULL buffer");
    FwiRefArgs a;
    float tab[FWI_NOALLOC_TABN];
    int rc = fwi_hip_internal_op_pml_args(g, &a, tab);  // rejects bad grids (npml 2..64, ...)
    if (rc) return rc;
    float *dtab = nullptr;
    rc = scratch_ptr(&dtab);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int n = 6 * (2 * g->npml + 1);
    for (int o = 0; o < n; o += FWI_NOALLOC_CHUNK) {
        TabChunk c;
        const int m = (n - o < FWI_NOALLOC_CHUNK) ? n - o : FWI_NOALLOC_CHUNK;
        for (int i = 0; i < FWI_NOALLOC_CHUNK; i++) c.v[i] = (i < m) ? tab[o + i] : 0.f;
        hipLaunchKernelGGL(k_tab_chunk, dim3(1), dim3(FWI_NOALLOC_CHUNK), 0, st, c, m, dtab + o);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fwi_hip_internal_err(FWI_HIP_EHIP, "op_pml_noalloc: table launch failed");
    }
    a.p = p; a.u = u; a.v = v; a.ux = ux; a.vz = vz;
    a.u0 = u0; a.u1 = u1; a.v1 = v1; a.ux0 = ux0; a.ux1 = ux1; a.vz1 = vz1;
    a.tab = dtab;
    for (int ph = 0; ph < 12; ph++) {
        const hipError_t e = fwi_launch_ref_phase(&a, ph, st);
        if (e != hipSuccess) return fwi_hip_internal_err(FWI_HIP_EHIP, "op_pml_noalloc: phase launch failed");
    }
    return FWI_HIP_OK;
}
