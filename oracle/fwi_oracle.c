/*
 * fwi_oracle.c -- CPU restatement of the wave-equation core of
 * i5fchen/FWI2DAcoustic-MPI-LBFGS, used ONLY as test infrastructure.
 *
 *   This file is the oracle: tests/, __graft_entry__.smoke() and the
 *   cpu_baseline leg of bench.py are the only code allowed to load it.
 *   The product (the HIP library) never links, loads or calls it.
 *
 * Built twice by oracle/Makefile: REAL=float  -> liboracle_f32.so (the FP32
 * oracle, arithmetic follows the reference expression by expression,
 * including the double promotions that C applies to its double literals)
 * and REAL=double -> liboracle_f64.so (the same algorithm in FP64).
 * Compiled with -ffp-contract=off so that every expression rounds as
 * written.
 *
 * Reference kernels restated (file:line in the reference repository):
 *   orc_cpml_coe        fwi_2d.c:110-137  (+ mirror fwi_2d.c:275-283)
 *   orc_fdcoe           fwi_2d.c:302-316
 *   orc_model_ext       fwi_2d.c:67-107
 *   orc_op_pml          fwi_2d.c:761-864  (a_lap_opt_2d_pml, free surface)
 *   orc_op_interior     fwi_2d.c:722-759  (a_lap_opt_2d)
 *   orc_time_integral   fwi_2d.c:866-880
 *   orc_write_bdr       fwi_2d.c:882-910  (strip layout of bc_ud / bc_lr)
 *   orc_read_bdr        fwi_2d.c:940-957
 *
 * The per-shot driver grad4_1shot is declared at header.h:283 and called at
 * MBotOpt.c:55,99 but has no body in the reference.  orc_forward /
 * orc_backward implement the driver this project defines for it (the same
 * definition the HIP path implements):
 *
 *   forward, it = 0..nt-1 (p^0 = p^-1 = 0, all CPML memory zero):
 *     a_lap_opt_2d_pml(p^it) -> ux, vz
 *     ux[src] += wav[it]*iarea                    (source term f^it)
 *     p^{it+1} = 2 p^it - p^{it-1} + ekp*(ux+vz)  (time_integral)
 *     syn[r][it] = p^{it+1}[rec r]
 *     save the 3-cell frame of the physical grid of p^{it+1} (write_bdr) and
 *     of X^it = ux+vz (the operator output, needed by the imaging condition)
 *
 *   misfit (host, FP64): J = 1/2 sum (syn - obs)^2, res = syn - obs
 *
 *   backward, n = nt-1..0 (mu = ekp * lambda, lambda = dJ/dp):
 *     mu^{n+1} = 2 mu^{n+2} - mu^{n+3} + ekp*(A^T(mu^{n+2}) + res[.][n] at rec)
 *       where A^T is the exact transpose of a_lap_opt_2d_pml over the time
 *       history (CPML memory run backward in time, orc_op_pml_adj)
 *     X^n   = a_lap_opt_2d(p^n) (+ f^n at the source) on cells >= 3 from the
 *             physical-grid edge, saved X^n on the 3-cell frame
 *     gK   += mu^{n+1} * X^n                   (zero-lag imaging, FP64 sum)
 *     p^{n-1} = 2 p^n - p^{n+1} + ekp*X^n, frame of p^{n-1} from the strips
 *   grad_v = 2 * gK / v   (ekp = v*v*dt*dt, so dJ/dv = dJ/dekp * 2 v dt^2
 *                          and dJ/dekp = gK / ekp)
 *
 * Parity pins for every function are in tests/test_oracle.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

/* Mpar.h:1 and fwi_sys.h:1 both define PI; fwi_2d.c includes fwi_sys.h after
 * Mpar.h, so the value the reference compiles with is this one. */
#define ORC_PI 3.141592654
#define ORC_MM 2 /* fwi_sys.h: MM (4th order => 2*MM-1 = 3 saved layers) */
#define NB (2 * ORC_MM - 1)

/* ------------------------------------------------------------------ */
/* a5: CPML coefficient table, fwi_2d.c:110-137 (npower = pml_base.dim = 2,
 * set at fwi_2d.c:257).  a,b have 2*npml+1 entries.                      */
void orc_cpml_coe(real *a0, real *b0, int npml, real dx, real dt,
                  real maxF, real maxV, real coefR)
{
    const real ndx = dx * 0.5;
    int npower = 2;
    real *d = (real *)calloc(2 * npml + 1, sizeof(real));
    real *alpha = (real *)calloc(2 * npml + 1, sizeof(real));
    real thickness_pml = npml * dx, PIf0 = ORC_PI * maxF, slope = PIf0 / thickness_pml;
    real d0 = -(npower + 1) * maxV * log(coefR) / (2.0 * thickness_pml);
    for (int i = 0; i < 2 * npml + 1; i++) {
        d[i] = (1.0 - i * ndx / thickness_pml);
        d[i] *= d[i];
        d[i] *= d0;
        alpha[i] = slope * (i * ndx);
        b0[i] = exp(-(d[i] + alpha[i]) * dt);
        a0[i] = d[i] / (d[i] + alpha[i]) * (b0[i] - 1.0);
    }
    free(d);
    free(alpha);
}

/* mirror table, fwi_2d.c:275-283 */
void orc_cpml_mirror(const real *a0, real *a1, int npml)
{
    for (int i = 0; i < 2 * npml + 1; i++) a1[i] = a0[2 * npml - i];
}

/* a6: FD coefficients, fwi_2d.c:302-316.  out = {c0x,c1x,ix,c0z,c1z,iz,
 * iz840,iz24} in the member order of fwi_fdcoe (header.h:101-107). */
void orc_fdcoe(real dx, real dz, real *out)
{
    real ix = 1.0 / dx, iz = 1.0 / dz;
    out[0] = 0.1129042e1 * ix;
    out[1] = -0.4301412e-1 * ix;
    out[2] = ix;
    out[3] = 0.1129042e1 * iz;
    out[4] = -0.4301412e-1 * iz;
    out[5] = iz;
    out[6] = iz / 840.0;
    out[7] = iz / 24.0;
}

/* a9: model extension under a free surface, fwi_2d.c:67-107.
 * model[n2][n1] -> emodel[en2][en1] (n1 = x, fastest). */
void orc_model_ext(const real *model, int n1, int n2, real *emodel, int en1, int en2)
{
    int gap1 = (en1 - n1) / 2, gap2 = (en2 - n2);
    for (int i2 = 0; i2 < en2 - gap2; i2++) {
        for (int i1 = gap1; i1 < en1 - gap1; i1++) emodel[i2 * en1 + i1] = model[i2 * n1 + i1 - gap1];
        for (int i1 = 0; i1 < gap1; i1++) emodel[i2 * en1 + i1] = model[i2 * n1 + 0];
        for (int i1 = en1 - gap1; i1 < en1; i1++) emodel[i2 * en1 + i1] = model[i2 * n1 + n1 - 1];
    }
    for (int i2 = en2 - gap2; i2 < en2; i2++)
        for (int i1 = gap1; i1 < en1 - gap1; i1++) emodel[i2 * en1 + i1] = model[(n2 - 1) * n1 + i1 - gap1];
    for (int i2 = en2 - gap2; i2 < en2; i2++)
        for (int i1 = en1 - gap1; i1 < en1; i1++) emodel[i2 * en1 + i1] = model[(n2 - 1) * n1 + n1 - 1];
    for (int i2 = en2 - gap2; i2 < en2; i2++)
        for (int i1 = 0; i1 < gap1; i1++) emodel[i2 * en1 + i1] = model[(n2 - 1) * n1 + 0];
}

/* ------------------------------------------------------------------ */
/* CPML tables used by the operator: tab = 6 rows of 2*npml+1:
 *   0: a0x  1: b0x  2: a1x  3: b1x  (x: left table and its mirror)
 *   4: a1z  5: b1z                   (z: bottom = mirror of the z table)
 * The top z tables (a0[1], b0[1]) are unused under a free surface. */
#define TAB(k, i) tab[(k) * (2 * npml + 1) + (i)]

/* a1: a_lap_opt_2d_pml, fwi_2d.c:761-864 (free surface, pnz = nz + npml).
 * p[pnz][pnx]; u[pnz][pnx+1]; v[pnz+1][pnx]; ux, vz[pnz][pnx].
 * psi: u0,u1[unz][npml+1]; v1[npml+1][vnx]; ux0,ux1[pnz][npml]; vz1[npml][pnx].
 * u, v must persist between calls exactly like the reference's (their
 * never-written entries u[:,0], u[:,unx-1], v[vnz-1] stay 0).            */
void orc_op_pml(const real *p, real *u, real *v, real *ux, real *vz,
                const real *coe, const real *tab, int npml, int pnx, int pnz,
                real *u0, real *u1, real *v1, real *ux0, real *ux1, real *vz1)
{
    const int unx = pnx + 1, unz = pnz, vnx = pnx, vnz = pnz + 1;
    const real c0x = coe[0], c1x = coe[1], ix_ = coe[2], c0z = coe[3], c1z = coe[4], iz_ = coe[5];
    const real iz840 = coe[6], iz24 = coe[7];
    const int n1 = npml + 1;
#define P(z, x) p[(z) * pnx + (x)]
#define U(z, x) u[(z) * unx + (x)]
#define V(z, x) v[(z) * vnx + (x)]
#define UX(z, x) ux[(z) * pnx + (x)]
#define VZ(z, x) vz[(z) * pnx + (x)]
#pragma omp parallel for
    for (int iz = 0; iz < unz; iz++)
        for (int ix = 2; ix < unx - 2; ix++)
            U(iz, ix) = c0x * (P(iz, ix) - P(iz, ix - 1)) + c1x * (P(iz, ix + 1) - P(iz, ix - 2));
#pragma omp parallel for
    for (int iz = 0; iz < unz; iz++) {
        U(iz, 1) = ix_ * (P(iz, 1) - P(iz, 0));
        U(iz, unx - 2) = ix_ * (P(iz, pnx - 1) - P(iz, pnx - 2));
    }
#pragma omp parallel for
    for (int iz = 0; iz < unz; iz++)
        for (int ix = 0; ix < npml + 1; ix++) {
            u0[iz * n1 + ix] = TAB(1, 2 * ix) * u0[iz * n1 + ix] + TAB(0, 2 * ix) * U(iz, ix);
            U(iz, ix) += u0[iz * n1 + ix];
            u1[iz * n1 + ix] = TAB(3, 2 * ix) * u1[iz * n1 + ix] + TAB(2, 2 * ix) * U(iz, ix + unx - npml - 1);
            U(iz, ix + unx - npml - 1) += u1[iz * n1 + ix];
        }
#pragma omp parallel for
    for (int iz = 2; iz < vnz - 2; iz++)
        for (int ix = 0; ix < vnx; ix++)
            V(iz, ix) = c0z * (P(iz, ix) - P(iz - 1, ix)) + c1z * (P(iz + 1, ix) - P(iz - 2, ix));
#pragma omp parallel for
    for (int ix = 0; ix < vnx; ix++) {
        V(0, ix) = (3675.0 * P(0, ix) - 1225.0 * P(1, ix) + 441.0 * P(2, ix) - 75.0 * P(3, ix)) * iz840;
        V(1, ix) = (-1085.0 * P(0, ix) + 1015.0 * P(1, ix) - 63.0 * P(2, ix) + 5.0 * P(3, ix)) * iz840;
        V(vnz - 2, ix) = (P(pnz - 1, ix) - P(pnz - 2, ix)) * iz_;
    }
#pragma omp parallel for
    for (int iz = 0; iz < npml + 1; iz++)
        for (int ix = 0; ix < vnx; ix++) {
            v1[iz * vnx + ix] = TAB(5, 2 * iz) * v1[iz * vnx + ix] + TAB(4, 2 * iz) * V(iz + vnz - npml - 1, ix);
            V(iz + vnz - npml - 1, ix) += v1[iz * vnx + ix];
        }
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++)
        for (int ix = 1; ix < pnx - 1; ix++)
            UX(iz, ix) = c0x * (U(iz, ix + 1) - U(iz, ix)) + c1x * (U(iz, ix + 2) - U(iz, ix - 1));
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++) {
        UX(iz, 0) = (U(iz, 1) - U(iz, 0)) * ix_;
        UX(iz, pnx - 1) = (U(iz, unx - 1) - U(iz, unx - 2)) * ix_;
    }
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++)
        for (int ix = 0; ix < npml; ix++) {
            ux0[iz * npml + ix] = TAB(1, 2 * ix + 1) * ux0[iz * npml + ix] + TAB(0, 2 * ix + 1) * UX(iz, ix);
            UX(iz, ix) += ux0[iz * npml + ix];
            ux1[iz * npml + ix] = TAB(3, 2 * ix + 1) * ux1[iz * npml + ix] + TAB(2, 2 * ix + 1) * UX(iz, ix + pnx - npml);
            UX(iz, ix + pnx - npml) += ux1[iz * npml + ix];
        }
#pragma omp parallel for
    for (int ix = 0; ix < pnx; ix++) {
        VZ(0, ix) = (-22. * V(0, ix) + 17. * V(1, ix) + 9. * V(2, ix) - 5. * V(3, ix) + V(4, ix)) * iz24;
        VZ(pnz - 1, ix) = (V(vnz - 1, ix) - V(vnz - 2, ix)) * iz_;
    }
#pragma omp parallel for
    for (int iz = 1; iz < pnz - 1; iz++)
        for (int ix = 0; ix < pnx; ix++)
            VZ(iz, ix) = c0z * (V(iz + 1, ix) - V(iz, ix)) + c1z * (V(iz + 2, ix) - V(iz - 1, ix));
#pragma omp parallel for
    for (int iz = 0; iz < npml; iz++)
        for (int ix = 0; ix < pnx; ix++) {
            vz1[iz * pnx + ix] = TAB(5, 2 * iz + 1) * vz1[iz * pnx + ix] + TAB(4, 2 * iz + 1) * VZ(iz + pnz - npml, ix);
            VZ(iz + pnz - npml, ix) += vz1[iz * pnx + ix];
        }
#undef P
#undef U
#undef V
#undef UX
#undef VZ
}

/* a3: a_lap_opt_2d, fwi_2d.c:722-759, on a grid p[pnz][pnx] without PML.
 * Entries outside the loop ranges are left untouched (as in the reference). */
void orc_op_interior(const real *p, real *u, real *v, real *ux, real *vz,
                     const real *coe, int pnx, int pnz)
{
    const int unx = pnx + 1, unz = pnz, vnx = pnx, vnz = pnz + 1;
    const real c0x = coe[0], c1x = coe[1], c0z = coe[3], c1z = coe[4];
#pragma omp parallel for
    for (int iz = 2; iz < unz - 2; iz++)
        for (int ix = 2; ix < unx - 2; ix++)
            u[iz * unx + ix] = c0x * (p[iz * pnx + ix] - p[iz * pnx + ix - 1]) + c1x * (p[iz * pnx + ix + 1] - p[iz * pnx + ix - 2]);
#pragma omp parallel for
    for (int iz = 2; iz < vnz - 2; iz++)
        for (int ix = 2; ix < vnx - 2; ix++)
            v[iz * vnx + ix] = c0z * (p[iz * pnx + ix] - p[(iz - 1) * pnx + ix]) + c1z * (p[(iz + 1) * pnx + ix] - p[(iz - 2) * pnx + ix]);
#pragma omp parallel for
    for (int iz = 3; iz < pnz - 3; iz++)
        for (int ix = 3; ix < pnx - 3; ix++)
            ux[iz * pnx + ix] = c0x * (u[iz * unx + ix + 1] - u[iz * unx + ix]) + c1x * (u[iz * unx + ix + 2] - u[iz * unx + ix - 1]);
#pragma omp parallel for
    for (int iz = 3; iz < pnz - 3; iz++)
        for (int ix = 3; ix < pnx - 3; ix++)
            vz[iz * pnx + ix] = c0z * (v[(iz + 1) * vnx + ix] - v[iz * vnx + ix]) + c1z * (v[(iz + 2) * vnx + ix] - v[(iz - 1) * vnx + ix]);
}

/* a2: time_integral, fwi_2d.c:866-880 */
void orc_time_integral(real *pnew, const real *pnow, const real *pold,
                       const real *pxx, const real *pzz, const real *ekp, int m, int n)
{
#pragma omp parallel for
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            pnew[i * n + j] = 2.0 * pnow[i * n + j] - pold[i * n + j] + ekp[i * n + j] * (pxx[i * n + j] + pzz[i * n + j]);
}

/* a4: one step of write_bdr, fwi_2d.c:891-910, without the disk chunking.
 * pf: f-grid field [nz+npml][nx+2npml]; ud[6][nx]; lr[nz][6].            */
void orc_write_bdr(real *ud, real *lr, const real *pf, int nx, int nz, int npml)
{
    const int pnx_f = nx + 2 * npml;
#pragma omp parallel for
    for (int i = 0; i < NB; i++)
        for (int j = 0; j < nx; j++) {
            ud[i * nx + j] = pf[i * pnx_f + j + npml];
            ud[(i + NB) * nx + j] = pf[(nz - i - 1) * pnx_f + j + npml];
        }
#pragma omp parallel for
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < NB; j++) {
            lr[i * 2 * NB + j] = pf[i * pnx_f + j + npml];
            lr[i * 2 * NB + j + NB] = pf[i * pnx_f + nx + npml - j - 1];
        }
}

/* a4: one step of read_bdr, fwi_2d.c:940-957, into a b-grid field pb[nz][nx] */
void orc_read_bdr(const real *ud, const real *lr, real *pb, int nx, int nz)
{
#pragma omp parallel for
    for (int i = 0; i < NB; i++)
        for (int j = 0; j < nx; j++) {
            pb[i * nx + j] = ud[i * nx + j];
            pb[(nz - i - 1) * nx + j] = ud[(i + NB) * nx + j];
        }
#pragma omp parallel for
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < NB; j++) {
            pb[i * nx + j] = lr[i * 2 * NB + j];
            pb[i * nx + nx - j - 1] = lr[i * 2 * NB + j + NB];
        }
}

/* ------------------------------------------------------------------ */
/* Exact transpose of orc_op_pml over the time history.
 * Forward, per axis: out = C_d( D2( C_h( D1 p ))) where D1 = the first
 * staggered derivative (p -> u or v nodes), C_h the CPML recursion on the
 * half nodes (psi = b psi + a x; y = x + psi), D2 the second derivative
 * (u/v nodes -> p nodes) and C_d the recursion on the p nodes.  Its
 * transpose over space and time is D1^T C_h^T D2^T C_d^T, where C^T is the
 * same recursion run backward in time (chi = b chi + a x; y = x + chi).
 * D1^T and D2^T gather, per node, the terms every row of the forward
 * definitions above scatters to it, in the order a row-by-row scatter
 * would add them, so they are the transposes by construction.
 * w[pnz][pnx] in; out[pnz][pnx] out; chi_* are the backward CPML states
 * (same shapes as the forward psi); tu[pnz][pnx+1], tv[pnz+1][pnx],
 * tw[pnz][pnx] are work arrays.                                          */
/* memcpy of an n-element field split over the threads */
static void par_copy(real *dst, const real *src, long n)
{
#pragma omp parallel for
    for (long b = 0; b < n; b += 4096) memcpy(dst + b, src + b, sizeof(real) * (n - b < 4096 ? n - b : 4096));
}
void orc_op_pml_adj(const real *w, real *out, const real *coe, const real *tab,
                    int npml, int pnx, int pnz,
                    real *cu0, real *cu1, real *cv1, real *cux0, real *cux1, real *cvz1,
                    real *tu, real *tv, real *tw)
{
    const int unx = pnx + 1, vnx = pnx, vnz = pnz + 1;
    const real c0x = coe[0], c1x = coe[1], ix_ = coe[2], c0z = coe[3], c1z = coe[4], iz_ = coe[5];
    const real iz840 = coe[6], iz24 = coe[7];
    const int n1 = npml + 1;

    /* ---------------- x chain ----------------
     * Every node is written by one thread with its terms in a fixed order,
     * so the result does not depend on the thread count.                 */
    par_copy(tw, w, (long)pnx * pnz);
    /* C_d^T on the ux nodes (transpose of fwi_2d.c:834-843) */
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++)
        for (int ix = 0; ix < npml; ix++) {
            cux0[iz * npml + ix] = TAB(1, 2 * ix + 1) * cux0[iz * npml + ix] + TAB(0, 2 * ix + 1) * tw[iz * pnx + ix];
            tw[iz * pnx + ix] += cux0[iz * npml + ix];
            cux1[iz * npml + ix] = TAB(3, 2 * ix + 1) * cux1[iz * npml + ix] + TAB(2, 2 * ix + 1) * tw[iz * pnx + ix + pnx - npml];
            tw[iz * pnx + ix + pnx - npml] += cux1[iz * npml + ix];
        }
    /* D2^T of the ux rows (fwi_2d.c:821-832).  Row ix of D2 writes
     * ux[ix] from u[ix-1..ix+2]; its transpose, gathered per u node j, adds
     * the terms of rows j-2, j-1, j, j+1 in that order, then the 2nd-order
     * edge rows 0 and pnx-1.                                             */
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++) {
        real *t = tu + iz * unx;
        const real *g = tw + iz * pnx;
        for (int j = 3; j <= pnx - 3; j++) {
            real acc = 0;
            acc += c1x * g[j - 2];
            acc += c0x * g[j - 1];
            acc -= c0x * g[j];
            acc -= c1x * g[j + 1];
            t[j] = acc;
        }
        for (int j = 0; j < unx; j++) {
            if (j == 3) j = pnx - 2; /* the nodes above */
            real acc = 0;
            if (j - 2 >= 1 && j - 2 <= pnx - 2) acc += c1x * g[j - 2];
            if (j - 1 >= 1 && j - 1 <= pnx - 2) acc += c0x * g[j - 1];
            if (j >= 1 && j <= pnx - 2) acc -= c0x * g[j];
            if (j + 1 >= 1 && j + 1 <= pnx - 2) acc -= c1x * g[j + 1];
            if (j == 1) acc += ix_ * g[0];
            if (j == 0) acc -= ix_ * g[0];
            if (j == unx - 1) acc += ix_ * g[pnx - 1];
            if (j == unx - 2) acc -= ix_ * g[pnx - 1];
            t[j] = acc;
        }
    }
    /* C_h^T on the u nodes (transpose of fwi_2d.c:786-795) */
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++)
        for (int ix = 0; ix < npml + 1; ix++) {
            cu0[iz * n1 + ix] = TAB(1, 2 * ix) * cu0[iz * n1 + ix] + TAB(0, 2 * ix) * tu[iz * unx + ix];
            tu[iz * unx + ix] += cu0[iz * n1 + ix];
            cu1[iz * n1 + ix] = TAB(3, 2 * ix) * cu1[iz * n1 + ix] + TAB(2, 2 * ix) * tu[iz * unx + ix + unx - npml - 1];
            tu[iz * unx + ix + unx - npml - 1] += cu1[iz * n1 + ix];
        }
    /* D1^T of the u rows (fwi_2d.c:775-784; u nodes 0 and unx-1 are zero):
     * row ix of D1 writes u[ix] from p[ix-2..ix+1]; gathered per p node j,
     * the terms of rows j-1, j, j+1, j+2 in that order, then the 2nd-order
     * rows 1 and unx-2.                                                  */
#pragma omp parallel for
    for (int iz = 0; iz < pnz; iz++) {
        real *o = out + iz * pnx;
        const real *t = tu + iz * unx;
        for (int j = 3; j <= pnx - 4; j++) {
            real acc = 0;
            acc += c1x * t[j - 1];
            acc += c0x * t[j];
            acc -= c0x * t[j + 1];
            acc -= c1x * t[j + 2];
            o[j] = acc;
        }
        for (int j = 0; j < pnx; j++) {
            if (j == 3) j = pnx - 3; /* the nodes above */
            real acc = 0;
            if (j - 1 >= 2 && j - 1 <= unx - 3) acc += c1x * t[j - 1];
            if (j >= 2 && j <= unx - 3) acc += c0x * t[j];
            if (j + 1 >= 2 && j + 1 <= unx - 3) acc -= c0x * t[j + 1];
            if (j + 2 >= 2 && j + 2 <= unx - 3) acc -= c1x * t[j + 2];
            if (j == 1) acc += ix_ * t[1];
            if (j == 0) acc -= ix_ * t[1];
            if (j == pnx - 1) acc += ix_ * t[unx - 2];
            if (j == pnx - 2) acc -= ix_ * t[unx - 2];
            o[j] = acc;
        }
    }

    /* ---------------- z chain ---------------- */
    par_copy(tw, w, (long)pnx * pnz);
    /* C_d^T on the vz nodes (transpose of fwi_2d.c:857-862) */
#pragma omp parallel for
    for (int iz = 0; iz < npml; iz++)
        for (int ix = 0; ix < pnx; ix++) {
            cvz1[iz * pnx + ix] = TAB(5, 2 * iz + 1) * cvz1[iz * pnx + ix] + TAB(4, 2 * iz + 1) * tw[(iz + pnz - npml) * pnx + ix];
            tw[(iz + pnz - npml) * pnx + ix] += cvz1[iz * pnx + ix];
        }
    /* D2^T of the vz rows (fwi_2d.c:846-855): the free-surface row 0 reads
     * v nodes 0..4 and the last row v nodes vnz-2, vnz-1; row iz in
     * [1, pnz-2] reads v nodes iz-1..iz+2.  Gathered per v node r: the
     * terms of rows 0 and pnz-1 first, then those of rows r-2, r-1, r, r+1. */
    {
        const double w0[5] = {-22., 17., 9., -5., 1.};
#pragma omp parallel for
        for (int r = 0; r < vnz; r++) {
            real *t = tv + r * vnx;
            if (r >= 5 && r <= pnz - 3) {
                const real *ga = tw + (r - 2) * pnx, *gb = ga + pnx, *gc = gb + pnx, *gd = gc + pnx;
                for (int ix = 0; ix < pnx; ix++) {
                    real acc = 0;
                    acc += c1z * ga[ix];
                    acc += c0z * gb[ix];
                    acc -= c0z * gc[ix];
                    acc -= c1z * gd[ix];
                    t[ix] = acc;
                }
                continue;
            }
            for (int ix = 0; ix < pnx; ix++) {
                real acc = 0;
                if (r <= 4) acc += (w0[r] * iz24) * tw[ix];
                if (r == vnz - 1) acc += iz_ * tw[(pnz - 1) * pnx + ix];
                if (r == vnz - 2) acc -= iz_ * tw[(pnz - 1) * pnx + ix];
                if (r - 2 >= 1 && r - 2 <= pnz - 2) acc += c1z * tw[(r - 2) * pnx + ix];
                if (r - 1 >= 1 && r - 1 <= pnz - 2) acc += c0z * tw[(r - 1) * pnx + ix];
                if (r >= 1 && r <= pnz - 2) acc -= c0z * tw[r * pnx + ix];
                if (r + 1 >= 1 && r + 1 <= pnz - 2) acc -= c1z * tw[(r + 1) * pnx + ix];
                t[ix] = acc;
            }
        }
    }
    /* C_h^T on the v nodes (transpose of fwi_2d.c:812-819) */
#pragma omp parallel for
    for (int iz = 0; iz < npml + 1; iz++)
        for (int ix = 0; ix < vnx; ix++) {
            cv1[iz * vnx + ix] = TAB(5, 2 * iz) * cv1[iz * vnx + ix] + TAB(4, 2 * iz) * tv[(iz + vnz - npml - 1) * vnx + ix];
            tv[(iz + vnz - npml - 1) * vnx + ix] += cv1[iz * vnx + ix];
        }
    /* D1^T of the v rows (fwi_2d.c:797-810; v node vnz-1 is zero): v rows
     * 0 and 1 (free surface) read p rows 0..3, row vnz-2 rows pnz-2 and
     * pnz-1, row iz in [2, vnz-3] rows iz-2..iz+1.  Gathered per p row r
     * and added to the x chain's result: the free-surface and last-row
     * terms first, then those of v rows r-1, r, r+1, r+2.                */
    {
        const double fs0[4] = {3675.0, -1225.0, 441.0, -75.0}, fs1[4] = {-1085.0, 1015.0, -63.0, 5.0};
#pragma omp parallel for
        for (int r = 0; r < pnz; r++) {
            real *o = out + r * pnx;
            if (r >= 4 && r <= pnz - 4) {
                const real *ta = tv + (r - 1) * vnx, *tb = ta + vnx, *tc = tb + vnx, *td = tc + vnx;
                for (int ix = 0; ix < pnx; ix++) {
                    real acc = o[ix];
                    acc += c1z * ta[ix];
                    acc += c0z * tb[ix];
                    acc -= c0z * tc[ix];
                    acc -= c1z * td[ix];
                    o[ix] = acc;
                }
                continue;
            }
            for (int ix = 0; ix < pnx; ix++) {
                real acc = o[ix];
                if (r <= 3) acc += (fs0[r] * iz840) * tv[ix] + (fs1[r] * iz840) * tv[vnx + ix];
                if (r == pnz - 1) acc += iz_ * tv[(vnz - 2) * vnx + ix];
                if (r == pnz - 2) acc -= iz_ * tv[(vnz - 2) * vnx + ix];
                if (r - 1 >= 2 && r - 1 <= vnz - 3) acc += c1z * tv[(r - 1) * vnx + ix];
                if (r >= 2 && r <= vnz - 3) acc += c0z * tv[r * vnx + ix];
                if (r + 1 >= 2 && r + 1 <= vnz - 3) acc -= c0z * tv[(r + 1) * vnx + ix];
                if (r + 2 >= 2 && r + 2 <= vnz - 3) acc -= c1z * tv[(r + 2) * vnx + ix];
                o[ix] = acc;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Driver (the body the reference's grad4_1shot never received).        */

typedef struct {
    int nx, nz, npml, pnx, pnz;
    real coe[8];
    real *tab;
} orc_geom;

static void geom_init(orc_geom *g, int nx, int nz, int npml, real dx, real dz, real dt,
                      real maxF, real maxV, real coefR)
{
    g->nx = nx;
    g->nz = nz;
    g->npml = npml;
    g->pnx = nx + 2 * npml; /* fwi_2d.c:155 */
    g->pnz = nz + npml;     /* fwi_2d.c:155, free surface */
    orc_fdcoe(dx, dz, g->coe);
    const int nt_ = 2 * npml + 1;
    g->tab = (real *)calloc(6 * nt_, sizeof(real));
    real *az = (real *)calloc(2 * nt_, sizeof(real));
    orc_cpml_coe(g->tab + 0 * nt_, g->tab + 1 * nt_, npml, dx, dt, maxF, maxV, coefR);
    orc_cpml_mirror(g->tab + 0 * nt_, g->tab + 2 * nt_, npml);
    orc_cpml_mirror(g->tab + 1 * nt_, g->tab + 3 * nt_, npml);
    orc_cpml_coe(az, az + nt_, npml, dz, dt, maxF, maxV, coefR);
    orc_cpml_mirror(az, g->tab + 4 * nt_, npml);
    orc_cpml_mirror(az + nt_, g->tab + 5 * nt_, npml);
    free(az);
}

/* CPML tables in the layout orc_op_pml takes (for tests). */
void orc_cpml_tables(int npml, real dx, real dz, real dt, real maxF, real maxV, real coefR, real *tab)
{
    orc_geom g;
    geom_init(&g, 8, 8, npml, dx, dz, dt, maxF, maxV, coefR);
    memcpy(tab, g.tab, sizeof(real) * 6 * (2 * npml + 1));
    free(g.tab);
}

/* ekp on the f-grid from a b-grid velocity: kp = v*v*dt*dt, then model_ext */
void orc_ekp_from_vel(const real *vel, int nx, int nz, int npml, real dt, real *ekp)
{
    real *kp = (real *)malloc(sizeof(real) * nx * nz);
    for (int i = 0; i < nx * nz; i++) kp[i] = vel[i] * vel[i] * dt * dt;
    orc_model_ext(kp, nx, nz, ekp, nx + 2 * npml, nz + npml);
    free(kp);
}

/* frame of the physical grid: strip[6*nx + 6*nz] = ud[6][nx] then lr[nz][6] */
static void frame_save_f(real *strip, const real *pf, const orc_geom *g)
{
    orc_write_bdr(strip, strip + 2 * NB * g->nx, pf, g->nx, g->nz, g->npml);
}

typedef struct {
    real *pold, *pnow, *pnew, *u, *v, *ux, *vz;
    real *u0, *u1, *v1, *ux0, *ux1, *vz1;
} orc_state;

static void state_alloc(orc_state *s, const orc_geom *g)
{
    const int pnx = g->pnx, pnz = g->pnz, n1 = g->npml + 1;
    s->pold = calloc((size_t)pnx * pnz, sizeof(real));
    s->pnow = calloc((size_t)pnx * pnz, sizeof(real));
    s->pnew = calloc((size_t)pnx * pnz, sizeof(real));
    s->u = calloc((size_t)(pnx + 1) * pnz, sizeof(real));
    s->v = calloc((size_t)pnx * (pnz + 1), sizeof(real));
    s->ux = calloc((size_t)pnx * pnz, sizeof(real));
    s->vz = calloc((size_t)pnx * pnz, sizeof(real));
    s->u0 = calloc((size_t)pnz * n1, sizeof(real));
    s->u1 = calloc((size_t)pnz * n1, sizeof(real));
    s->v1 = calloc((size_t)n1 * pnx, sizeof(real));
    s->ux0 = calloc((size_t)pnz * g->npml, sizeof(real));
    s->ux1 = calloc((size_t)pnz * g->npml, sizeof(real));
    s->vz1 = calloc((size_t)g->npml * pnx, sizeof(real));
}

static void state_free(orc_state *s)
{
    free(s->pold); free(s->pnow); free(s->pnew); free(s->u); free(s->v); free(s->ux); free(s->vz);
    free(s->u0); free(s->u1); free(s->v1); free(s->ux0); free(s->ux1); free(s->vz1);
}

/* Forward propagation of one shot.
 *   vel[nz][nx]: physical velocity (m/s); isx,isz: source cell (physical
 *   grid, fwi_2d.c:363-364); wav[nt]: wavelet (already scaled);
 *   irx,irz[nr]: receiver cells (physical grid, fwi_2d.c:366-371).
 * Outputs (each may be NULL):
 *   syn[nr][nt]                       receiver traces
 *   pstrip[nt][6nx+6nz]               frame of p^{it+1} (write_bdr layout)
 *   xstrip[nt][6nx+6nz]               frame of X^{it} = ux + vz (+ source)
 *   plast[2][nz][nx]                  p^{nt}, p^{nt-1} on the physical grid
 *   xall[nt][nz][nx]                  X^{it} on the whole physical grid
 * Returns 0.                                                             */
int orc_forward(int nx, int nz, int npml, real dx, real dz, real dt,
                real maxF, real maxV, real coefR, const real *vel,
                int nt, int isx, int isz, const real *wav,
                int nr, const int *irx, const int *irz,
                real *syn, real *pstrip, real *xstrip, real *plast, real *xall)
{
    orc_geom g;
    geom_init(&g, nx, nz, npml, dx, dz, dt, maxF, maxV, coefR);
    const int pnx = g.pnx, pnz = g.pnz;
    real *ekp = (real *)malloc(sizeof(real) * pnx * pnz);
    orc_ekp_from_vel(vel, nx, nz, npml, dt, ekp);
    const real iarea = 1.0 / (dx * dz); /* fwi_2d.c:216 */
    const size_t sl = (size_t)(6 * nx + 6 * nz);
    orc_state s;
    state_alloc(&s, &g);
    real *xf = (real *)malloc(sizeof(real) * pnx * pnz);
    for (int it = 0; it < nt; it++) {
        orc_op_pml(s.pnow, s.u, s.v, s.ux, s.vz, g.coe, g.tab, npml, pnx, pnz,
                   s.u0, s.u1, s.v1, s.ux0, s.ux1, s.vz1);
        s.ux[isz * pnx + isx + npml] += wav[it] * iarea;
        orc_time_integral(s.pnew, s.pnow, s.pold, s.ux, s.vz, ekp, pnz, pnx);
        for (int r = 0; r < nr && syn; r++) syn[(size_t)r * nt + it] = s.pnew[irz[r] * pnx + irx[r] + npml];
        if (pstrip) frame_save_f(pstrip + (size_t)it * sl, s.pnew, &g);
        if (xstrip || xall) {
#pragma omp parallel for
            for (int i = 0; i < pnx * pnz; i++) xf[i] = s.ux[i] + s.vz[i];
            if (xstrip) frame_save_f(xstrip + (size_t)it * sl, xf, &g);
            if (xall)
                for (int z = 0; z < nz; z++)
                    memcpy(xall + ((size_t)it * nz + z) * nx, xf + z * pnx + npml, sizeof(real) * nx);
        }
        real *t = s.pold;
        s.pold = s.pnow;
        s.pnow = s.pnew;
        s.pnew = t;
    }
    if (plast)
        for (int z = 0; z < nz; z++) {
            memcpy(plast + z * nx, s.pnow + z * pnx + npml, sizeof(real) * nx);
            memcpy(plast + (size_t)nx * nz + z * nx, s.pold + z * pnx + npml, sizeof(real) * nx);
        }
    state_free(&s);
    free(xf);
    free(ekp);
    free(g.tab);
    return 0;
}

/* L2 waveform misfit (msf_pool.h:1 declares L2norm without a body):
 * res = syn - obs, returns 1/2 sum res^2 accumulated in double. */
double orc_misfit_l2(const real *syn, const real *obs, real *res, long n)
{
    double J = 0.0;
    for (long i = 0; i < n; i++) {
        real r = syn[i] - obs[i];
        res[i] = r;
        J += 0.5 * (double)r * (double)r;
    }
    return J;
}

/* Backward sweep of one shot: adjoint propagation of res (exact transpose
 * of the forward), source reconstruction from the saved frames, zero-lag
 * imaging.
 *   res[nr][nt]: adjoint source (dJ/dsyn).
 *   pstrip, xstrip, plast: from orc_forward (used when xall == NULL).
 *   xall[nt][nz][nx]: if non-NULL, the stored forward X^n is used for the
 *     imaging instead of the reconstruction (reference answer for tests).
 * Outputs:
 *   gk[nz][nx] (double): sum_n mu^{n+1} X^n   (= ekp * dJ/dekp)
 *   grad[nz][nx] (double, may be NULL): dJ/dv = 2 gk / v
 *   musrc[nt] (may be NULL): mu^{n+1} at the source cell, n = 0..nt-1
 *     (sum_n musrc[n]*wav[n]*iarea = <syn, res> : the dot-product test)  */
int orc_backward(int nx, int nz, int npml, real dx, real dz, real dt,
                 real maxF, real maxV, real coefR, const real *vel,
                 int nt, int isx, int isz, const real *wav,
                 int nr, const int *irx, const int *irz, const real *res,
                 const real *pstrip, const real *xstrip, const real *plast,
                 const real *xall, double *gk, double *grad, real *musrc)
{
    orc_geom g;
    geom_init(&g, nx, nz, npml, dx, dz, dt, maxF, maxV, coefR);
    const int pnx = g.pnx, pnz = g.pnz;
    real *ekp = (real *)malloc(sizeof(real) * pnx * pnz);
    orc_ekp_from_vel(vel, nx, nz, npml, dt, ekp);
    const real iarea = 1.0 / (dx * dz);
    const size_t sl = (size_t)(6 * nx + 6 * nz);
    const size_t nf = (size_t)pnx * pnz, nb = (size_t)nx * nz;
    const int n1 = npml + 1;

    /* adjoint state on the f-grid */
    real *m1 = calloc(nf, sizeof(real)); /* mu^{n+2} */
    real *m2 = calloc(nf, sizeof(real)); /* mu^{n+3}, overwritten by mu^{n+1} */
    real *aout = calloc(nf, sizeof(real));
    real *cu0 = calloc((size_t)pnz * n1, sizeof(real)), *cu1 = calloc((size_t)pnz * n1, sizeof(real));
    real *cv1 = calloc((size_t)n1 * pnx, sizeof(real));
    real *cux0 = calloc((size_t)pnz * npml, sizeof(real)), *cux1 = calloc((size_t)pnz * npml, sizeof(real));
    real *cvz1 = calloc((size_t)npml * pnx, sizeof(real));
    real *tu = calloc((size_t)pnz * (pnx + 1), sizeof(real)), *tv = calloc((size_t)(pnz + 1) * pnx, sizeof(real));
    real *tw = calloc(nf, sizeof(real));

    /* reconstruction state on the b-grid */
    real *rn = calloc(nb, sizeof(real)), *rp = calloc(nb, sizeof(real)), *rm = calloc(nb, sizeof(real));
    real *bu = calloc((size_t)(nx + 1) * nz, sizeof(real)), *bv = calloc((size_t)nx * (nz + 1), sizeof(real));
    real *bux = calloc(nb, sizeof(real)), *bvz = calloc(nb, sizeof(real));
    real *xb = calloc(nb, sizeof(real)), *ekpb = calloc(nb, sizeof(real));
    real *fr = calloc(sl, sizeof(real));
    for (int z = 0; z < nz; z++) memcpy(ekpb + z * nx, ekp + z * pnx + npml, sizeof(real) * nx);
    if (!xall) {
        memcpy(rn, plast, sizeof(real) * nb);      /* p^{nt}   */
        memcpy(rp, plast + nb, sizeof(real) * nb); /* p^{nt-1} */
    }
    /* after the swap at the top of the first iteration: rn = p^{nt-1}, rp = p^{nt} */
    { real *t = rn; rn = rp; rp = t; }
    memset(gk, 0, sizeof(double) * nb);

    for (int n = nt - 1; n >= 0; n--) {
        /* ---- adjoint: mu^{n+1} = 2 mu^{n+2} - mu^{n+3} + ekp (A^T mu^{n+2} + R) */
        orc_op_pml_adj(m1, aout, g.coe, g.tab, npml, pnx, pnz, cu0, cu1, cv1, cux0, cux1, cvz1, tu, tv, tw);
        for (int r = 0; r < nr; r++) aout[irz[r] * pnx + irx[r] + npml] += res[(size_t)r * nt + n];
#pragma omp parallel for
        for (size_t i = 0; i < nf; i++) m2[i] = 2.0 * m1[i] - m2[i] + ekp[i] * aout[i];
        { real *t = m1; m1 = m2; m2 = t; } /* m1 = mu^{n+1}, m2 = mu^{n+2} */
        if (musrc) musrc[n] = m1[isz * pnx + isx + npml];

        /* ---- X^n on the physical grid */
        if (xall) {
            memcpy(xb, xall + (size_t)n * nb, sizeof(real) * nb);
        } else {
            orc_op_interior(rn, bu, bv, bux, bvz, g.coe, nx, nz);
#pragma omp parallel for
            for (int z = 3; z < nz - 3; z++)
                for (int x = 3; x < nx - 3; x++) xb[z * nx + x] = bux[z * nx + x];
            if (isz >= 3 && isz < nz - 3 && isx >= 3 && isx < nx - 3) xb[isz * nx + isx] += wav[n] * iarea;
#pragma omp parallel for
            for (int z = 3; z < nz - 3; z++)
                for (int x = 3; x < nx - 3; x++) xb[z * nx + x] = xb[z * nx + x] + bvz[z * nx + x];
            orc_read_bdr(xstrip + (size_t)n * sl, xstrip + (size_t)n * sl + 2 * NB * nx, xb, nx, nz);
        }
        /* ---- zero-lag imaging */
#pragma omp parallel for
        for (int z = 0; z < nz; z++)
            for (int x = 0; x < nx; x++)
                gk[z * nx + x] += (double)m1[z * pnx + x + npml] * (double)xb[z * nx + x];
        /* ---- reconstruction p^{n-1} = 2 p^n - p^{n+1} + ekp X^n */
        if (!xall && n >= 1) {
#pragma omp parallel for
            for (size_t i = 0; i < nb; i++) rm[i] = 2.0 * rn[i] - rp[i] + ekpb[i] * xb[i];
            if (n >= 2) orc_read_bdr(pstrip + (size_t)(n - 2) * sl, pstrip + (size_t)(n - 2) * sl + 2 * NB * nx, rm, nx, nz);
            else {
                memset(fr, 0, sizeof(real) * sl);
                orc_read_bdr(fr, fr + 2 * NB * nx, rm, nx, nz);
            }
            real *t = rp;
            rp = rn;
            rn = rm;
            rm = t;
        }
    }
    if (grad)
        for (size_t i = 0; i < nb; i++) grad[i] = 2.0 * gk[i] / (double)vel[i];

    free(m1); free(m2); free(aout); free(cu0); free(cu1); free(cv1); free(cux0); free(cux1); free(cvz1);
    free(tu); free(tv); free(tw); free(rn); free(rp); free(rm); free(bu); free(bv); free(bux); free(bvz);
    free(xb); free(ekpb); free(fr); free(ekp); free(g.tab);
    return 0;
}

/* The full per-shot gradient as the HIP path computes it: forward with
 * frame saving, L2 misfit, backward with reconstruction.  Returns J.     */
double orc_gradient(int nx, int nz, int npml, real dx, real dz, real dt,
                    real maxF, real maxV, real coefR, const real *vel,
                    int nt, int isx, int isz, const real *wav,
                    int nr, const int *irx, const int *irz, const real *obs,
                    real *syn, double *grad, double *gk)
{
    const size_t sl = (size_t)(6 * nx + 6 * nz);
    real *ps = malloc(sizeof(real) * sl * nt), *xs = malloc(sizeof(real) * sl * nt);
    real *pl = malloc(sizeof(real) * 2 * nx * nz), *res = malloc(sizeof(real) * nr * nt);
    orc_forward(nx, nz, npml, dx, dz, dt, maxF, maxV, coefR, vel, nt, isx, isz, wav, nr, irx, irz,
                syn, ps, xs, pl, NULL);
    double J = orc_misfit_l2(syn, obs, res, (long)nr * nt);
    orc_backward(nx, nz, npml, dx, dz, dt, maxF, maxV, coefR, vel, nt, isx, isz, wav, nr, irx, irz,
                 res, ps, xs, pl, NULL, gk, grad, NULL);
    free(ps); free(xs); free(pl); free(res);
    return J;
}

int orc_real_size(void) { return (int)sizeof(real); }
