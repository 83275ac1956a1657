/*
 * TEST INFRASTRUCTURE ONLY - the CPU oracle for the extrapolate-and-image path.
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * and the cpu_baseline leg of bench.py may load the library built from it.
 * The HIP library (imaging2d_wlsq_amd/csrc) never links or calls it.
 *
 * It restates the reference CPU path in plain C (no MKL, no thrust):
 *
 *   oracle_extrapolate        cpp_src/extrapolation_cpu.cpp:11-107
 *                             (argument order of cpp_src/interface.py:50-58)
 *       - padded wavefields nf x (nx+2M) per shot with ZERO halos, as the
 *         reference intends (py_src/wavefieldPropagation/extrapolation.py:48-51,
 *         unittests/test_py_src_extrapolation.py:135-144; cpp_src/wfPad.h:89-115
 *         forgets to zero them because its halo loop runs over My = 0)
 *       - per (l, j, s, i): new[j][i+M] = sum_k op[l][j][i][k] * old[j][i+k],
 *         an UNCONJUGATED dot (cpp_src/mkl_linalg.cpp:7-11, cblas_cdotu_sub),
 *         summed in tap order k = 0..L-1 from a zero accumulator
 *       - ping-pong of old/new (extrapolation_cpu.cpp:62-65)
 *       - imaging after the step: conv = sum_j new_f * conj(new_b) in FP32,
 *         sequential over j (extrapolation_cpu.cpp:69-85); image[s][l][i] = Re(conv)
 *       - in-place write-back of the final wavefields with shot stride nt*nx
 *         (extrapolation_cpu.cpp:91-98)
 *       - 64-bit offsets everywhere (the reference computes l*nf*nx*L in int,
 *         extrapolation_cpu.cpp:40)
 *
 *   oracle_extrapolate_ktable the same step, but the per-point operator is
 *                             gathered from a compact (nk x L) table by
 *                             q = oracle_kappa_index(omega_j, slowness[l][i]),
 *                             the lookup of py_src/operatorsPreparation/
 *                             preparation.py:109 + util.py:5-8 (nearest kappa,
 *                             first index on ties) evaluated in FP32.
 *
 *   oracle_expand_table       builds the reference-layout table
 *                             (nextrap, nf, nx, L) from a compact table with
 *                             the same lookup (preparation.py:101-115 with the
 *                             symmetric operator for every x, main.py:66).
 *
 * Complex arithmetic follows thrust::complex<float> (cpp_src/types.h:4-9):
 * a*b = (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re), no fused multiply-add
 * (built with -ffp-contract=off).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float re, im;
} c32;

static inline c32 cmul(c32 a, c32 b)
{
    c32 r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

static inline c32 cadd(c32 a, c32 b)
{
    c32 r;
    r.re = a.re + b.re;
    r.im = a.im + b.im;
    return r;
}

static inline c32 cconj(c32 a)
{
    c32 r;
    r.re = a.re;
    r.im = -a.im;
    return r;
}

/*
 * Nearest kappa sample to omega*slowness on the uniform grid
 * kappa_q = kappa0 + q*dkappa, q in [0, nk); ties go to the lower index, as
 * numpy's argmin does in util.py:5-8. Evaluated in FP32 with one rounding per
 * operation (no fused multiply-add) so that every implementation, NumPy
 * float32 included, gets the same q:
 *   t = omega * slowness; t = t - kappa0; u = t * inv_dkappa;
 *   u = clamp(u, 0, nk-1); q = ceil(u - 0.5)
 * inv_dkappa is (float)(1.0 / (double)dkappa), computed by the caller.
 */
int oracle_kappa_index(float omega, float slowness, float kappa0, float inv_dkappa, int nk)
{
    float t = omega * slowness;
    t = t - kappa0;
    float u = t * inv_dkappa;
    float hi = (float)(nk - 1);
    if (!(u > 0.0f))
        u = 0.0f; /* also maps NaN to 0 */
    if (u > hi)
        u = hi;
    float v = u - 0.5f;
    return (int)ceilf(v);
}

float oracle_inv_dkappa(float dkappa)
{
    return (float)(1.0 / (double)dkappa);
}

/* one dot of length L, tap order, zero start (cblas_cdotu_sub semantics) */
static inline c32 cdotu(const c32 *w, const c32 *x, int L)
{
    c32 acc = {0.0f, 0.0f};
    for (int k = 0; k < L; ++k)
        acc = cadd(acc, cmul(w[k], x[k]));
    return acc;
}

/*
 * Shared driver for both table kinds. When `table_f` is non-NULL the operator
 * for (l, j, i) is table_f + ((l*nf + j)*nx + i)*L (reference layout);
 * otherwise it is ktab_f + q*L with q from the kappa lookup.
 */
static int run(int ns, int nextrap, int nz, int nx, int nf, int nt, int M,
               const c32 *table_f, const c32 *table_b,
               const float *slowness, const float *omega, float kappa0, float dkappa, int nk,
               const c32 *ktab_f, const c32 *ktab_b,
               c32 *wave, c32 *shot, float *image)
{
    if (ns < 0 || nextrap < 0 || nz < nextrap || nx <= 0 || nf < 0 || nt < nf || M < 0)
        return 1;
    const int L = 2 * M + 1;
    const size_t dim_x = (size_t)nx + 2 * (size_t)M;
    const size_t pad = (size_t)nf * dim_x;
    const size_t nsz = (size_t)ns;

    c32 *of = (c32 *)calloc(nsz * pad + 1, sizeof(c32));
    c32 *nf_ = (c32 *)calloc(nsz * pad + 1, sizeof(c32));
    c32 *ob = (c32 *)calloc(nsz * pad + 1, sizeof(c32));
    c32 *nb = (c32 *)calloc(nsz * pad + 1, sizeof(c32));
    if (!of || !nf_ || !ob || !nb) {
        free(of); free(nf_); free(ob); free(nb);
        return 2;
    }
    for (size_t s = 0; s < nsz; ++s)
        for (size_t j = 0; j < (size_t)nf; ++j)
            for (size_t i = 0; i < (size_t)nx; ++i) {
                of[s * pad + j * dim_x + i + M] = wave[s * (size_t)nt * nx + j * nx + i];
                ob[s * pad + j * dim_x + i + M] = shot[s * (size_t)nt * nx + j * nx + i];
            }

    const float inv_dk = ktab_f ? oracle_inv_dkappa(dkappa) : 0.0f;

    for (int l = 0; l < nextrap; ++l) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 0; j < nf; ++j) {
            for (size_t s = 0; s < nsz; ++s) {
                for (int i = 0; i < nx; ++i) {
                    const c32 *wf, *wb;
                    if (table_f) {
                        size_t o = (((size_t)l * nf + j) * nx + i) * (size_t)L;
                        wf = table_f + o;
                        wb = table_b + o;
                    } else {
                        int q = oracle_kappa_index(omega[j], slowness[(size_t)l * nx + i],
                                                   kappa0, inv_dk, nk);
                        wf = ktab_f + (size_t)q * L;
                        wb = ktab_b + (size_t)q * L;
                    }
                    const size_t row = s * pad + (size_t)j * dim_x;
                    nf_[row + i + M] = cdotu(wf, of + row + i, L);
                    nb[row + i + M] = cdotu(wb, ob + row + i, L);
                }
            }
        }
        /* swap (extrapolation_cpu.cpp:62-65) */
        c32 *t;
        t = of; of = nf_; nf_ = t;
        t = ob; ob = nb; nb = t;

        /* imaging (extrapolation_cpu.cpp:69-85): sequential over j per (s, i) */
#pragma omp parallel for schedule(dynamic, 1)
        for (int s = 0; s < ns; ++s) {
            for (int i = 0; i < nx; ++i) {
                c32 conv = {0.0f, 0.0f};
                for (int j = 0; j < nf; ++j) {
                    size_t o = (size_t)s * pad + (size_t)j * dim_x + i + M;
                    conv = cadd(conv, cmul(of[o], cconj(ob[o])));
                }
                image[(size_t)s * nz * nx + (size_t)l * nx + i] = conv.re;
            }
        }
    }

    /* write-back (extrapolation_cpu.cpp:91-98), shot stride nt*nx */
    for (size_t s = 0; s < nsz; ++s)
        for (size_t j = 0; j < (size_t)nf; ++j)
            for (size_t i = 0; i < (size_t)nx; ++i) {
                wave[s * (size_t)nt * nx + j * nx + i] = of[s * pad + j * dim_x + i + M];
                shot[s * (size_t)nt * nx + j * nx + i] = ob[s * pad + j * dim_x + i + M];
            }
    free(of); free(nf_); free(ob); free(nb);
    return 0;
}

/* cpp_src/extrapolation_cpu.cpp:11-14 argument order */
int oracle_extrapolate(int ns, int nextrap, int nz, int nx, int nf, int nt, int M,
                       const c32 *w_op_fs_forw, const c32 *w_op_fs_back,
                       c32 *wave_fs, c32 *shot_fs, float *image)
{
    return run(ns, nextrap, nz, nx, nf, nt, M, w_op_fs_forw, w_op_fs_back,
               NULL, NULL, 0.0f, 1.0f, 0, NULL, NULL, wave_fs, shot_fs, image);
}

/* compact-table form; slowness is (>= nextrap, nx) row-major, omega has nf entries */
int oracle_extrapolate_ktable(int ns, int nextrap, int nz, int nx, int nf, int nt, int M,
                              const float *slowness, const float *omega,
                              float kappa0, float dkappa, int nk,
                              const c32 *table_forw, const c32 *table_back,
                              c32 *wave_fs, c32 *shot_fs, float *image)
{
    if (nk <= 0 || !(dkappa > 0.0f))
        return 1;
    return run(ns, nextrap, nz, nx, nf, nt, M, NULL, NULL, slowness, omega, kappa0, dkappa, nk,
               table_forw, table_back, wave_fs, shot_fs, image);
}

/* reference-layout expansion: out[l][j][i][:] = table[q(l,j,i)][:] */
int oracle_expand_table(int nextrap, int nf, int nx, int M,
                        const float *slowness, const float *omega,
                        float kappa0, float dkappa, int nk,
                        const c32 *table, c32 *out)
{
    if (nk <= 0 || !(dkappa > 0.0f))
        return 1;
    const int L = 2 * M + 1;
    const float inv_dk = oracle_inv_dkappa(dkappa);
#pragma omp parallel for schedule(static)
    for (int l = 0; l < nextrap; ++l)
        for (int j = 0; j < nf; ++j)
            for (int i = 0; i < nx; ++i) {
                int q = oracle_kappa_index(omega[j], slowness[(size_t)l * nx + i], kappa0, inv_dk, nk);
                memcpy(out + (((size_t)l * nf + j) * nx + i) * (size_t)L,
                       table + (size_t)q * L, sizeof(c32) * (size_t)L);
            }
    return 0;
}
