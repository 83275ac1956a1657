// gfx950 kernels of the WLSQ extrapolate-and-image path.
//
// Semantics (one shot, one depth step l, frequency j, position i, L = 2M+1):
//   new[j][i] = sum_k W[l][j][i][k] * old[j][i+k-M]   (old = 0 outside [0, nx))
//   image[l][i] += Re( new_f[j][i] * conj(new_b[j][i]) ) summed over j
// as in py_src/wavefieldPropagation/extrapolation.py:57-64 and
// cpp_src/extrapolation_cpu.cpp:39-89. The dot is unconjugated
// (cpp_src/mkl_linalg.cpp:7-11) and taps are summed in order k = 0..L-1.
//
// Two paths:
//  * plain: the per-point operators come from an expanded table in HBM
//    (tap-major, one coalesced load per tap); one launch per depth step
//    updates both fields and writes per-block imaging partials. HBM-bound.
//  * fused (k-table): a workgroup owns complete rows (S shots x W frequencies)
//    of one shot group in LDS for ALL depth steps, gathers each operator from
//    a compact (nk x L) table by the kappa lookup and adds the image with
//    atomics. FP32-VALU bound.
#include "wlsq_internal.h"

namespace wlsq {
namespace {

// acc + a*b (complex); the product's parts are formed with fused multiply-adds.
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b)
{
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}

// Re(a * conj(b))
__device__ __forceinline__ float re_mul_conj(float2 a, float2 b)
{
    return fmaf(a.y, b.y, a.x * b.x);
}

// Nearest kappa sample to omega*s on kappa_q = kappa0 + q/inv_dk, ties to the
// lower q (numpy argmin in py_src/operatorsPreparation/util.py:5-8). One
// rounding per operation and no contraction, so the host-side expansion and
// the CPU restatement compute the same q bit for bit.
__device__ __forceinline__ int kappa_index(float omega, float s, float kappa0, float inv_dk, int nk)
{
#pragma clang fp contract(off)
    float t = omega * s;
    t = t - kappa0;
    float u = t * inv_dk;
    const float hi = (float)(nk - 1);
    u = (u > 0.0f) ? u : 0.0f;
    u = (u > hi) ? hi : u;
    return (int)ceilf(u - 0.5f);
}

// Position p of a padded row lives at sub-row (p mod R), slot p / R: lanes
// that read or write p = R*lane + c for a fixed c touch consecutive slots.
template <int R>
__device__ __forceinline__ int swz(int p, int Q)
{
    return (p & (R - 1)) * Q + p / R;
}

// ------------------------------------------------------------------ fused kernel
template <int L_, int R, int S, bool CONJ, bool SYM, bool PAIRED, int C, int MAXT>
__global__ void __launch_bounds__(MAXT) fused_kernel(const FusedArgs a)
{
    constexpr int LOG2R = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : R == 2 ? 1 : 0;
    static_assert((1 << LOG2R) == R, "R must be a power of two");
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    const Item it = a.items[blockIdx.x];
    const Group *gp = a.groups + it.group;
    const int nshots = gp->nshots;
    const int nx = gp->nx;
    const int nz = gp->nz;
    const int ld = gp->ld;
    const int L = L_ > 0 ? L_ : a.L;
    const int M = (L - 1) >> 1;
    const int nstrips = (nx + R - 1) / R;
    const int nxr = nstrips * R;
    // padded row: M halo + nxr + M halo + 1 slot read by the final window slide
    const int Q = (nxr + 2 * M + R) / R;
    const int rowlen = Q * R;
    const int W = it.nw;
    const int nrows = S * W;
    if (W * nstrips > C * NT)
        return;  // the planner never builds such a task
    float2 *rowsF = lds;
    float2 *rowsB = lds + (size_t)nrows * rowlen;
    float *imgrow = reinterpret_cast<float *>(lds + 2 * (size_t)nrows * rowlen);

    // Initial rows: interior from the fields, zero halos and padding.
    for (int e = tid; e < nrows * rowlen; e += NT) {
        const int row = e / rowlen;
        const int p = e - row * rowlen;
        const int s = row / W;
        const int w = row - s * W;
        const int x = p - M;
        float2 vf = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
        if (s < nshots && x >= 0 && x < nx) {
            const long long r = (long long)(it.j0 + w) * ld + x;
            vf = a.F[gp->f_off[s] + r];
            vb = a.B[gp->b_off[s] + r];
        }
        const int o = row * rowlen + swz<R>(p, Q);
        rowsF[o] = vf;
        rowsB[o] = vb;
    }
    const bool shared_img = gp->img_shared != 0;
    const int img_rows = shared_img ? 1 : nshots;  // image rows reduced in LDS
    if (a.img_lds)
        for (int e = tid; e < img_rows * nxr; e += NT)
            imgrow[e] = 0.f;

    long long ioff[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        ioff[s] = gp->img_off[s];
    const int img_ld = gp->img_ld;
    const int slow_ld = gp->slow_ld;
    const float *slow0 = a.slow + gp->slow_off;
    // tables have rows of Lp = L + 1 entries (16-byte aligned): taps are
    // gathered in pairs, one 16-byte load per pair and table
    const float2 *__restrict__ TF = a.TF;
    const float2 *__restrict__ TB = a.TB;
    const int Lp = L + 1;

    // Each thread owns C strips (task slots tid, tid + NT, ...): slot ->
    // (frequency row w, strip of R consecutive x).
    const size_t shot_rows = (size_t)W * rowlen;
    bool actv[C];
    int wv[C], x0v[C];
    float omv[C];
#pragma unroll
    for (int cs = 0; cs < C; ++cs) {
        const int slot = tid + cs * NT;
        actv[cs] = slot < W * nstrips;
        wv[cs] = actv[cs] ? slot / nstrips : 0;
        x0v[cs] = actv[cs] ? (slot - wv[cs] * nstrips) * R : 0;
        omv[cs] = actv[cs] ? a.omega[gp->omega_off + it.j0 + wv[cs]] : 0.f;
    }
    __syncthreads();

    float2 accF_[C][S][R], accB_[C][S][R];
    for (int l = 0; l < nz; ++l) {
#pragma unroll
        for (int cs = 0; cs < C; ++cs) {
        if (actv[cs]) {
            float2 (&accF)[S][R] = accF_[cs];
            float2 (&accB)[S][R] = accB_[cs];
            const int x0 = x0v[cs];
            const float om = omv[cs];
            const float2 *rF = rowsF + (size_t)wv[cs] * rowlen;  // shot s adds s*W*rowlen
            const float2 *rB = rowsB + (size_t)wv[cs] * rowlen;
            const float *srow = slow0 + (long long)l * slow_ld;
            int qo[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int xr = min(x0 + r, nx - 1);
                qo[r] = kappa_index(om, srow[xr], a.kappa0, a.inv_dk, a.nk) * Lp;
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    accF[s][r] = make_float2(0.f, 0.f);
                    accB[s][r] = make_float2(0.f, 0.f);
                }
            }
            // Row position p of this thread's rows lives at sub-row (p mod R),
            // slot p / R; x0 is a multiple of R, so p = x0 + j is at sub-row
            // (j mod R), slot xs + j / R.
            const int xs = x0 / R;
            auto ldF = [&](int s, int j) { return rF[s * shot_rows + (j & (R - 1)) * Q + xs + (j >> LOG2R)]; };
            auto ldB = [&](int s, int j) { return rB[s * shot_rows + (j & (R - 1)) * Q + xs + (j >> LOG2R)]; };
            if (!SYM) {
                // new[x] = sum_k T[k] old[x + k - M]: a window of R inputs slides
                // right by one per tap. Taps go in chunks of R so the window
                // registers are back in place after each chunk.
                float2 wF[S][R], wB[S][R];
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        wF[s][r] = ldF(s, r);
                        wB[s][r] = ldB(s, r);
                    }
                }
                // coefficients of taps k, k+1 (k even) for every x of the strip
                auto fetch2 = [&](const int k, float2 (&f0)[R], float2 (&f1)[R], float2 (&b0)[R],
                                  float2 (&b1)[R]) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const float4 v = *reinterpret_cast<const float4 *>(TF + qo[r] + k);
                        f0[r] = make_float2(v.x, v.y);
                        f1[r] = make_float2(v.z, v.w);
                        if (CONJ) {
                            b0[r] = make_float2(v.x, -v.y);
                            b1[r] = make_float2(v.z, -v.w);
                        } else {
                            const float4 u = *reinterpret_cast<const float4 *>(TB + qo[r] + k);
                            b0[r] = make_float2(u.x, u.y);
                            b1[r] = make_float2(u.z, u.w);
                        }
                    }
                };
                auto tap = [&](const int k, const float2 (&cf)[R], const float2 (&cb)[R]) {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            accF[s][r] = cmac(accF[s][r], cf[r], wF[s][r]);
                            accB[s][r] = cmac(accB[s][r], cb[r], wB[s][r]);
                        }
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r + 1 < R; ++r) {
                            wF[s][r] = wF[s][r + 1];
                            wB[s][r] = wB[s][r + 1];
                        }
                        wF[s][R - 1] = ldF(s, R + k);  // the slot past the last tap is padding
                        wB[s][R - 1] = ldB(s, R + k);
                    }
                };
                auto tap2 = [&](const int k, const bool second) {
                    float2 f0[R], f1[R], b0[R], b1[R];
                    fetch2(k, f0, f1, b0, b1);
                    tap(k, f0, b0);
                    if (second)
                        tap(k + 1, f1, b1);
                    // wide windows: keep the next pair's loads from being hoisted
                    // over this one (bounds live registers)
                    if constexpr (R >= 8)
                        __builtin_amdgcn_sched_barrier(0);
                };
                auto tap1 = [&](const int k) {
                    float2 f0[R], b0[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        f0[r] = TF[qo[r] + k];
                        b0[r] = CONJ ? make_float2(f0[r].x, -f0[r].y) : TB[qo[r] + k];
                    }
                    tap(k, f0, b0);
                    if constexpr (R >= 8)
                        __builtin_amdgcn_sched_barrier(0);
                };
                const int nch = L / R;  // R is even: chunks start at even taps
#pragma unroll 1
                for (int c = 0; c < nch; ++c) {
                    if constexpr (PAIRED) {
#pragma unroll
                        for (int r2 = 0; r2 < R; r2 += 2)
                            tap2(c * R + r2, true);
                    } else {
#pragma unroll
                        for (int r2 = 0; r2 < R; ++r2)
                            tap1(c * R + r2);
                    }
                }
                if constexpr (PAIRED) {
                    if (L_ > 0) {
#pragma unroll
                        for (int k = (L_ / R) * R; k < L_; k += 2)
                            tap2(k, k + 1 < L_);
                    } else {
#pragma unroll 1
                        for (int k = nch * R; k < L; k += 2)
                            tap2(k, k + 1 < L);
                    }
                } else {
                    if (L_ > 0) {
#pragma unroll
                        for (int k = (L_ / R) * R; k < L_; ++k)
                            tap1(k);
                    } else {
#pragma unroll 1
                        for (int k = nch * R; k < L; ++k)
                            tap1(k);
                    }
                }
            } else {
                // Tap-symmetric operator (T[k] == T[L-1-k]):
                //   new[x] = T[M] old[x] + sum_{k<M} T[k] (old[x+k-M] + old[x+M-k]).
                // A left window slides right and a right window slides left;
                // after the M pairs the left window holds old[x] for the centre.
                float2 lF[S][R], lB[S][R], rFw[S][R], rBw[S][R];
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        lF[s][r] = ldF(s, r);
                        lB[s][r] = ldB(s, r);
                        rFw[s][r] = ldF(s, r + 2 * M);
                        rBw[s][r] = ldB(s, r + 2 * M);
                    }
                }
                auto fetch2 = [&](const int k, float2 (&f0)[R], float2 (&f1)[R], float2 (&b0)[R],
                                  float2 (&b1)[R]) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const float4 v = *reinterpret_cast<const float4 *>(TF + qo[r] + k);
                        f0[r] = make_float2(v.x, v.y);
                        f1[r] = make_float2(v.z, v.w);
                        if (CONJ) {
                            b0[r] = make_float2(v.x, -v.y);
                            b1[r] = make_float2(v.z, -v.w);
                        } else {
                            const float4 u = *reinterpret_cast<const float4 *>(TB + qo[r] + k);
                            b0[r] = make_float2(u.x, u.y);
                            b1[r] = make_float2(u.z, u.w);
                        }
                    }
                };
                auto pair = [&](const int k, const float2 (&cf)[R], const float2 (&cb)[R]) {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const float2 sf = make_float2(lF[s][r].x + rFw[s][r].x, lF[s][r].y + rFw[s][r].y);
                            const float2 sb = make_float2(lB[s][r].x + rBw[s][r].x, lB[s][r].y + rBw[s][r].y);
                            accF[s][r] = cmac(accF[s][r], cf[r], sf);
                            accB[s][r] = cmac(accB[s][r], cb[r], sb);
                        }
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r + 1 < R; ++r) {
                            lF[s][r] = lF[s][r + 1];
                            lB[s][r] = lB[s][r + 1];
                        }
                        lF[s][R - 1] = ldF(s, R + k);
                        lB[s][R - 1] = ldB(s, R + k);
#pragma unroll
                        for (int r = R - 1; r > 0; --r) {
                            rFw[s][r] = rFw[s][r - 1];
                            rBw[s][r] = rBw[s][r - 1];
                        }
                        rFw[s][0] = ldF(s, 2 * M - 1 - k);
                        rBw[s][0] = ldB(s, 2 * M - 1 - k);
                    }
                };
                auto pair2 = [&](const int k, const bool second) {
                    float2 f0[R], f1[R], b0[R], b1[R];
                    fetch2(k, f0, f1, b0, b1);
                    pair(k, f0, b0);
                    if (second)
                        pair(k + 1, f1, b1);
                    if constexpr (R >= 8)
                        __builtin_amdgcn_sched_barrier(0);
                };
                auto pair1 = [&](const int k) {
                    float2 f0[R], b0[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        f0[r] = TF[qo[r] + k];
                        b0[r] = CONJ ? make_float2(f0[r].x, -f0[r].y) : TB[qo[r] + k];
                    }
                    pair(k, f0, b0);
                    if constexpr (R >= 8)
                        __builtin_amdgcn_sched_barrier(0);
                };
                const int nch = M / R;
#pragma unroll 1
                for (int c = 0; c < nch; ++c) {
                    if constexpr (PAIRED) {
#pragma unroll
                        for (int r2 = 0; r2 < R; r2 += 2)
                            pair2(c * R + r2, true);
                    } else {
#pragma unroll
                        for (int r2 = 0; r2 < R; ++r2)
                            pair1(c * R + r2);
                    }
                }
                constexpr int M_ = L_ > 0 ? (L_ - 1) / 2 : 0;
                if constexpr (PAIRED) {
                    if (L_ > 0) {
#pragma unroll
                        for (int k = (M_ / R) * R; k < M_; k += 2)
                            pair2(k, k + 1 < M_);
                    } else {
#pragma unroll 1
                        for (int k = nch * R; k < M; k += 2)
                            pair2(k, k + 1 < M);
                    }
                } else {
                    if (L_ > 0) {
#pragma unroll
                        for (int k = (M_ / R) * R; k < M_; ++k)
                            pair1(k);
                    } else {
#pragma unroll 1
                        for (int k = nch * R; k < M; ++k)
                            pair1(k);
                    }
                }
                float2 cf[R], cb[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cf[r] = TF[qo[r] + M];
                    cb[r] = CONJ ? make_float2(cf[r].x, -cf[r].y) : TB[qo[r] + M];
                }
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        accF[s][r] = cmac(accF[s][r], cf[r], lF[s][r]);
                        accB[s][r] = cmac(accB[s][r], cb[r], lB[s][r]);
                    }
                }
            }
            // imaging condition of this step. With img_lds the task's rows are
            // summed in LDS (over its frequencies, and over its shots when they
            // share an image) and flushed once per (depth, x) below.
            if (a.img_lds) {
                if (shared_img) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        float v = 0.f;
#pragma unroll
                        for (int s = 0; s < S; ++s)
                            if (s < nshots)
                                v += re_mul_conj(accF[s][r], accB[s][r]);
                        if (x0 + r < nx)
                            atomicAdd(&imgrow[x0 + r], v);
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        if (s < nshots) {
#pragma unroll
                            for (int r = 0; r < R; ++r)
                                if (x0 + r < nx)
                                    atomicAdd(&imgrow[s * nxr + x0 + r],
                                              re_mul_conj(accF[s][r], accB[s][r]));
                        }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s < nshots) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int x = x0 + r;
                            if (x < nx)
                                unsafeAtomicAdd(a.img + ioff[s] + (long long)l * img_ld + x,
                                                re_mul_conj(accF[s][r], accB[s][r]));
                        }
                    }
                }
            }
        }
        }
        __syncthreads();
#pragma unroll
        for (int cs = 0; cs < C; ++cs) {
            if (!actv[cs])
                continue;
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int x = x0v[cs] + r;
                    if (x < nx) {
                        const int o = swz<R>(x + M, Q);
                        rowsF[s * shot_rows + (size_t)wv[cs] * rowlen + o] = accF_[cs][s][r];
                        rowsB[s * shot_rows + (size_t)wv[cs] * rowlen + o] = accB_[cs][s][r];
                    }
                }
            }
        }
        if (a.img_lds) {
            // consecutive lanes flush consecutive x: one 32-byte sector of the
            // image carries 8 atomic adds
            for (int e = tid; e < img_rows * nx; e += NT) {
                const int s = e / nx;
                const int x = e - s * nx;
                const float v = imgrow[s * nxr + x];
                imgrow[s * nxr + x] = 0.f;
                unsafeAtomicAdd(a.img + gp->img_off[s] + (long long)l * img_ld + x, v);
            }
        }
        __syncthreads();
    }

    if (a.write_back) {
        for (int e = tid; e < nrows * nx; e += NT) {
            const int row = e / nx;
            const int x = e - row * nx;
            const int s = row / W;
            const int ww = row - s * W;
            if (s < nshots) {
                const int o = row * rowlen + swz<R>(x + M, Q);
                const long long r = (long long)(it.j0 + ww) * ld + x;
                a.F[gp->f_off[s] + r] = rowsF[o];
                a.B[gp->b_off[s] + r] = rowsB[o];
            }
        }
    }
}

// ------------------------------------------------------------------ plain path
constexpr int kPlainTX = 64;  // x per block
constexpr int kPlainTJ = 4;   // frequencies per block

// One depth step of both fields for S shots per thread, tap-major operators
// op[(j*L + k)*nx + x] of this depth (revOp layout, cpp_src/revOp.c:4-30).
// The old rows (+M halos, zero outside [0, nx)) of the block's 4 frequencies
// are staged in LDS; each tap is one coalesced 512-byte load per wavefront
// per field. The image partial of the block (its 4 frequencies, fixed order)
// goes to part[s][blockIdx.y][x].
template <int S, int L_>
__global__ void __launch_bounds__(256) plain_step_kernel(
    const float2 *__restrict__ opF, const float2 *__restrict__ opB,
    const float2 *__restrict__ oldF, const float2 *__restrict__ oldB, long long in_stride,
    float2 *__restrict__ newF, float2 *__restrict__ newB, float *__restrict__ part,
    long long part_stride, int nshots, int nf, int nx, int M, int Lrt)
{
    extern __shared__ float2 sm[];
    const int L = L_ > 0 ? L_ : Lrt;
    const int tx = threadIdx.x;
    const int ty = threadIdx.y;
    const int x = blockIdx.x * kPlainTX + tx;
    const int j = blockIdx.y * kPlainTJ + ty;
    const int s0 = blockIdx.z * S;
    const int TW = kPlainTX + 2 * M;
    const int tile = S * kPlainTJ * TW;
    float2 *smF = sm;
    float2 *smB = sm + tile;
    float *red = reinterpret_cast<float *>(sm + 2 * tile);

    for (int e = ty * kPlainTX + tx; e < tile; e += kPlainTX * kPlainTJ) {
        const int s = e / (kPlainTJ * TW);
        const int rem = e - s * (kPlainTJ * TW);
        const int jj = rem / TW;
        const int p = rem - jj * TW;
        const int xg = blockIdx.x * kPlainTX + p - M;
        const int jg = blockIdx.y * kPlainTJ + jj;
        float2 vf = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
        if (s0 + s < nshots && jg < nf && xg >= 0 && xg < nx) {
            const long long o = (long long)(s0 + s) * in_stride + (long long)jg * nx + xg;
            vf = oldF[o];
            vb = oldB[o];
        }
        smF[e] = vf;
        smB[e] = vb;
    }
    __syncthreads();

    float2 aF[S], aB[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        aF[s] = make_float2(0.f, 0.f);
        aB[s] = make_float2(0.f, 0.f);
    }
    const bool valid = x < nx && j < nf;
    if (valid) {
        const float2 *cF = opF + (long long)j * L * nx + x;
        const float2 *cB = opB + (long long)j * L * nx + x;
#pragma unroll 8
        for (int k = 0; k < L; ++k) {
            const float2 cf = cF[(long long)k * nx];
            const float2 cb = cB[(long long)k * nx];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int o = (s * kPlainTJ + ty) * TW + tx + k;
                aF[s] = cmac(aF[s], cf, smF[o]);
                aB[s] = cmac(aB[s], cb, smB[o]);
            }
        }
        const long long plane = (long long)nf * nx;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s0 + s < nshots) {
                const long long o = (long long)(s0 + s) * plane + (long long)j * nx + x;
                newF[o] = aF[s];
                newB[o] = aB[s];
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        red[(s * kPlainTJ + ty) * kPlainTX + tx] = valid ? re_mul_conj(aF[s], aB[s]) : 0.f;
    __syncthreads();
    if (ty == 0 && x < nx) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s0 + s < nshots) {
                float v = red[(s * kPlainTJ + 0) * kPlainTX + tx];
#pragma unroll
                for (int jj = 1; jj < kPlainTJ; ++jj)
                    v += red[(s * kPlainTJ + jj) * kPlainTX + tx];
                part[(long long)(s0 + s) * part_stride + (long long)blockIdx.y * nx + x] = v;
            }
        }
    }
}

// image[s][l][x] = sum over g (in order) of part[s][l][g][x]
__global__ void plain_reduce_kernel(const float *__restrict__ part, float *__restrict__ image,
                                    long long image_shot_stride, int nextrap, int groups, int nx)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int s = blockIdx.z;
    if (x >= nx)
        return;
    const float *p = part + ((long long)s * nextrap + l) * groups * (long long)nx + x;
    float v = 0.f;
    for (int g = 0; g < groups; ++g)
        v += p[(long long)g * nx];
    image[(long long)s * image_shot_stride + (long long)l * nx + x] = v;
}

// tap[(slice*L + k)*nx + x] = nat[(slice*nx + x)*L + k]
__global__ void transpose_tapmajor_kernel(const float2 *__restrict__ nat, float2 *__restrict__ tap,
                                          int nx, int L)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const long long slice = blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (x >= nx)
        return;
    const float2 *src = nat + (slice * nx + x) * L;
    float2 *dst = tap + slice * L * (long long)nx + x;
    for (int k = 0; k < L; ++k)
        dst[(long long)k * nx] = src[k];
}

__global__ void copy_rows_kernel(const float2 *__restrict__ src, long long src_stride,
                                 float2 *__restrict__ dst, long long dst_stride, int nrows, int nx)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const long long n = (long long)nrows * nx;
    if (e >= n)
        return;
    dst[(long long)s * dst_stride + e] = src[(long long)s * src_stride + e];
}

// out = table[q(l, j, x)] in natural (l, j, x, k) or tap-major (l, j, k, x) order
__global__ void expand_kernel(const float *__restrict__ slow, int slow_ld,
                              const float *__restrict__ omega, float kappa0, float inv_dk, int nk,
                              const float2 *__restrict__ table, float2 *__restrict__ out, int nf,
                              int nx, int L, int layout)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int l = blockIdx.z;
    if (x >= nx)
        return;
    const int q = kappa_index(omega[j], slow[(long long)l * slow_ld + x], kappa0, inv_dk, nk);
    const float2 *src = table + (long long)q * L;
    const long long slice = (long long)l * nf + j;
    if (layout == WLSQ_LAYOUT_TAPMAJOR) {
        float2 *dst = out + slice * L * (long long)nx + x;
        for (int k = 0; k < L; ++k)
            dst[(long long)k * nx] = src[k];
    } else {
        float2 *dst = out + (slice * nx + x) * L;
        for (int k = 0; k < L; ++k)
            dst[k] = src[k];
    }
}

// out[q][0..L] = in[q][0..L-1], 0: rows padded to an even length
__global__ void table_pad_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int nk,
                                 int L)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int Lp = L + 1;
    if (e >= (long long)nk * Lp)
        return;
    const long long q = e / Lp;
    const int k = (int)(e - q * Lp);
    out[e] = k < L ? in[q * L + k] : make_float2(0.f, 0.f);
}

__global__ void fill_f32_kernel(float *p, long long n, float v)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride)
        p[e] = v;
}

template <typename K>
hipError_t ensure_lds(K kernel, size_t bytes)
{
    if (bytes <= 64 * 1024)
        return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int L_, int R, int S, bool CONJ, bool SYM, bool PAIRED, int C, int MAXT>
hipError_t launch_fused_t(int n_items, int block, size_t lds, hipStream_t st, const FusedArgs &a)
{
    auto k = fused_kernel<L_, R, S, CONJ, SYM, PAIRED, C, MAXT>;
    hipError_t e = ensure_lds(k, lds);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k, dim3(n_items), dim3(block), lds, st, a);
    return hipGetLastError();
}

template <int L_, int R, int S, int C, int MAXT>
hipError_t launch_fused_RS(const FusedShape &sh, int n_items, size_t lds, hipStream_t st,
                           const FusedArgs &a)
{
    const int b = sh.block;
#define WLSQ_FUSED_P(P)                                                                          \
    if (sh.sym)                                                                                  \
        return sh.conj ? launch_fused_t<L_, R, S, true, true, P, C, MAXT>(n_items, b, lds, st, a) \
                       : launch_fused_t<L_, R, S, false, true, P, C, MAXT>(n_items, b, lds, st, a); \
    return sh.conj ? launch_fused_t<L_, R, S, true, false, P, C, MAXT>(n_items, b, lds, st, a)    \
                   : launch_fused_t<L_, R, S, false, false, P, C, MAXT>(n_items, b, lds, st, a);
    if (sh.paired) {
        WLSQ_FUSED_P(true)
    } else {
        WLSQ_FUSED_P(false)
    }
#undef WLSQ_FUSED_P
}

template <int L_>
hipError_t launch_fused_L(const FusedShape &sh, int n_items, size_t lds, hipStream_t st,
                          const FusedArgs &a)
{
    if (sh.C == 1) {
        if (sh.R == 4 && sh.S == 1)
            return launch_fused_RS<L_, 4, 1, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 4 && sh.S == 2)
            return launch_fused_RS<L_, 4, 2, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 4 && sh.S == 4)
            return launch_fused_RS<L_, 4, 4, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 8 && sh.S == 1)
            return launch_fused_RS<L_, 8, 1, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 2 && sh.S == 2)
            return launch_fused_RS<L_, 2, 2, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 2 && sh.S == 4)
            return launch_fused_RS<L_, 2, 4, 1, 512>(sh, n_items, lds, st, a);
    } else if (sh.C == 2) {
        if (sh.R == 8 && sh.S == 1)
            return launch_fused_RS<L_, 8, 1, 2, 512>(sh, n_items, lds, st, a);
    }
    return hipErrorInvalidValue;
}

}  // namespace

size_t fused_lds_bytes(int nx, int M, int R, int S, int W, int img_rows)
{
    const size_t nstrips = (size_t)(nx + R - 1) / R;
    const size_t nxr = nstrips * R;
    const size_t Q = (nxr + 2 * (size_t)M + R) / R;
    const size_t rowlen = Q * R;
    return 2 * (size_t)S * W * rowlen * sizeof(float2) + (size_t)img_rows * nxr * sizeof(float);
}

hipError_t launch_fused(const FusedShape &sh, int n_items, size_t lds_bytes, hipStream_t st,
                        const FusedArgs &a)
{
    if (n_items <= 0)
        return hipSuccess;
    if (sh.block <= 0 || sh.block > 512 || sh.block % 64 != 0)
        return hipErrorInvalidValue;
    switch (a.L) {
    case 25: return launch_fused_L<25>(sh, n_items, lds_bytes, st, a);
    case 31: return launch_fused_L<31>(sh, n_items, lds_bytes, st, a);
    case 51: return launch_fused_L<51>(sh, n_items, lds_bytes, st, a);
    default: return launch_fused_L<0>(sh, n_items, lds_bytes, st, a);
    }
}

template <int S, int L_>
hipError_t launch_plain_t(dim3 grid, dim3 block, size_t lds, hipStream_t st, const float2 *opF,
                          const float2 *opB, const float2 *oldF, const float2 *oldB,
                          long long in_stride, float2 *newF, float2 *newB, float *part,
                          long long part_stride, int nshots, int nf, int nx, int M, int L)
{
    auto k = plain_step_kernel<S, L_>;
    hipError_t e = ensure_lds(k, lds);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k, grid, block, lds, st, opF, opB, oldF, oldB, in_stride, newF, newB, part,
                       part_stride, nshots, nf, nx, M, L);
    return hipGetLastError();
}

template <int S>
hipError_t launch_plain_S(dim3 grid, dim3 block, size_t lds, hipStream_t st, const float2 *opF,
                          const float2 *opB, const float2 *oldF, const float2 *oldB,
                          long long in_stride, float2 *newF, float2 *newB, float *part,
                          long long part_stride, int nshots, int nf, int nx, int M)
{
    const int L = 2 * M + 1;
    switch (L) {
    case 25: return launch_plain_t<S, 25>(grid, block, lds, st, opF, opB, oldF, oldB, in_stride, newF, newB, part, part_stride, nshots, nf, nx, M, L);
    case 31: return launch_plain_t<S, 31>(grid, block, lds, st, opF, opB, oldF, oldB, in_stride, newF, newB, part, part_stride, nshots, nf, nx, M, L);
    case 51: return launch_plain_t<S, 51>(grid, block, lds, st, opF, opB, oldF, oldB, in_stride, newF, newB, part, part_stride, nshots, nf, nx, M, L);
    default: return launch_plain_t<S, 0>(grid, block, lds, st, opF, opB, oldF, oldB, in_stride, newF, newB, part, part_stride, nshots, nf, nx, M, L);
    }
}

hipError_t launch_plain_step(int S, const float2 *opF, const float2 *opB, const float2 *oldF,
                             const float2 *oldB, long long in_shot_stride, float2 *newF,
                             float2 *newB, float *part, long long part_stride, int nshots, int nf,
                             int nx, int M, hipStream_t st)
{
    if (nshots <= 0 || nf <= 0 || nx <= 0)
        return hipSuccess;
    const int G = plain_groups(nf);
    dim3 block(kPlainTX, kPlainTJ);
    dim3 grid((nx + kPlainTX - 1) / kPlainTX, G, (nshots + S - 1) / S);
    const size_t tile = (size_t)S * kPlainTJ * (kPlainTX + 2 * M);
    const size_t lds = 2 * tile * sizeof(float2) + (size_t)S * kPlainTJ * kPlainTX * sizeof(float);
    switch (S) {
    case 1: return launch_plain_S<1>(grid, block, lds, st, opF, opB, oldF, oldB, in_shot_stride, newF, newB, part, part_stride, nshots, nf, nx, M);
    case 2: return launch_plain_S<2>(grid, block, lds, st, opF, opB, oldF, oldB, in_shot_stride, newF, newB, part, part_stride, nshots, nf, nx, M);
    case 4: return launch_plain_S<4>(grid, block, lds, st, opF, opB, oldF, oldB, in_shot_stride, newF, newB, part, part_stride, nshots, nf, nx, M);
    default: return hipErrorInvalidValue;
    }
}

int plain_groups(int nf) { return (nf + kPlainTJ - 1) / kPlainTJ; }

hipError_t launch_plain_reduce(const float *part, float *image, int nshots,
                               long long image_shot_stride, int nextrap, int groups, int nx,
                               hipStream_t st)
{
    if (nshots <= 0 || nextrap <= 0 || nx <= 0)
        return hipSuccess;
    dim3 block(256);
    dim3 grid((nx + 255) / 256, nextrap, nshots);
    hipLaunchKernelGGL(plain_reduce_kernel, grid, block, 0, st, part, image, image_shot_stride,
                       nextrap, groups, nx);
    return hipGetLastError();
}

hipError_t launch_transpose_tapmajor(const float2 *nat, float2 *tap, long long n_slices, int nx,
                                     int L, hipStream_t st)
{
    if (n_slices <= 0 || nx <= 0)
        return hipSuccess;
    long long gy = n_slices, gz = 1;
    while (gy > 65535) {
        if (gy % 2)
            return hipErrorInvalidValue;  // caller splits odd large counts
        gy /= 2;
        gz *= 2;
    }
    dim3 block(256);
    dim3 grid((nx + 255) / 256, (unsigned)gy, (unsigned)gz);
    hipLaunchKernelGGL(transpose_tapmajor_kernel, grid, block, 0, st, nat, tap, nx, L);
    return hipGetLastError();
}

hipError_t launch_copy_rows(const float2 *src, long long src_shot_stride, float2 *dst,
                            long long dst_shot_stride, int nshots, int nrows, int nx,
                            hipStream_t st)
{
    const long long n = (long long)nrows * nx;
    if (nshots <= 0 || n <= 0)
        return hipSuccess;
    dim3 block(256);
    dim3 grid((unsigned)((n + 255) / 256), nshots);
    hipLaunchKernelGGL(copy_rows_kernel, grid, block, 0, st, src, src_shot_stride, dst,
                       dst_shot_stride, nrows, nx);
    return hipGetLastError();
}

hipError_t launch_expand(const float *slow, int slow_ld, const float *omega, float kappa0,
                         float inv_dk, int nk, const float2 *table, float2 *out, int nextrap,
                         int nf, int nx, int L, int layout, hipStream_t st)
{
    if (nextrap <= 0 || nf <= 0 || nx <= 0)
        return hipSuccess;
    dim3 block(256);
    dim3 grid((nx + 255) / 256, nf, nextrap);
    hipLaunchKernelGGL(expand_kernel, grid, block, 0, st, slow, slow_ld, omega, kappa0, inv_dk, nk,
                       table, out, nf, nx, L, layout);
    return hipGetLastError();
}

hipError_t launch_table_pad(const float2 *in, float2 *out, int nk, int L, hipStream_t st)
{
    const long long n = (long long)nk * (L + 1);
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(table_pad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out,
                       nk, L);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace wlsq
