// gfx950 fused extrapolate-and-image kernel (compact operator tables).
//
// A workgroup owns complete rows (S shots x W frequencies) of one shot group
// in LDS for all its depth steps, gathers each operator from a compact
// (nk x L) table by the kappa lookup and adds the image with atomics: from
// per-frequency partial rows in LDS, or, where those do not fit next to the
// field rows, straight from the threads (the XP variant of the one-shot
// 1024-thread shapes: lane-contiguous through a wave-private LDS transpose,
// into image planes that a reduction kernel adds to the image after the launch).
// Semantics as in py_src/wavefieldPropagation/extrapolation.py:57-64 and
// cpp_src/extrapolation_cpu.cpp:39-89: an unconjugated dot over the taps,
// summed in tap order, or, for tap-symmetric tables, as M tap pairs plus the
// centre tap (the same sum in a different rounding order).
//
// This file is compiled once per (WLSQ_L, WLSQ_PAIRED): taps 25, 31, 51 or 0
// (any L, runtime loop bounds) and 8- or 16-byte coefficient gathers, so the
// variants build in parallel. Each object defines launch_fused_L<L>_P<P>.
#include "wlsq_device.h"

#ifndef WLSQ_L
#error "compile with -DWLSQ_L=<taps or 0> -DWLSQ_PAIRED=<0|1>"
#endif

namespace wlsq {
namespace {

// R consecutive floats at p (p is 4R-byte aligned: x0 is a multiple of R)
template <int R>
__device__ __forceinline__ void store_row(float *p, const float (&v)[R])
{
    if constexpr (R == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int r = 0; r < R; r += 4)
            *reinterpret_cast<float4 *>(p + r) = make_float4(v[r], v[r + 1], v[r + 2], v[r + 3]);
    }
}

// Orders LDS accesses of the lanes of one wavefront around it (the hardware
// executes a wavefront's LDS instructions in order; this keeps the compiler
// from moving them across).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ fused kernel
// LDS holds one 16-byte element (F, B) per position of each padded row, so a
// window load is one ds_read_b128 for both fields. Position p of a row lives
// at sub-row (p mod R), slot p / R (swz): lanes that read p = R*lane + c for a
// fixed c touch consecutive 16-byte slots.
template <int L_, int R, int S, bool CONJ, bool SYM, bool PAIRED, int C, int MAXT, bool XP,
          bool EXACT, bool XACC>
__global__ void __launch_bounds__(MAXT) fused_kernel(const FusedArgs a)
{
    static_assert(!(SYM && EXACT), "the CPU path's rounding sums the taps in order");
    // image sums formed from the new rows (img_lds 2): compiled for one shot
    // per group (a third branch in the step's imaging code makes the
    // multi-shot shapes spill)
    constexpr bool ROWS_IMG = S == 1;
    // complex multiply-add and imaging term: fused multiply-adds, or (EXACT)
    // the CPU path's separately rounded products
    auto mac = [](float2 acc, float2 u, float2 v) {
        if constexpr (EXACT)
            return cmac_exact(acc, u, v);
        else
            return cmac(acc, u, v);
    };
    auto rmc = [](float2 u, float2 v) {
        if constexpr (EXACT)
            return re_mul_conj_exact(u, v);
        else
            return re_mul_conj(u, v);
    };
    constexpr int LOG2R = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : R == 2 ? 1 : 0;
    static_assert((1 << LOG2R) == R, "R must be a power of two");
    constexpr int M_ = L_ > 0 ? (L_ - 1) / 2 : 0;
    extern __shared__ float4 lds4[];
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    const Item it = a.items[blockIdx.x];
    const Group *gp = a.groups + it.group;
    const int nshots = gp->nshots;
    const int nx = gp->nx;
    const int nz = gp->nz;
    const int ld = gp->ld;
    const int L = L_ > 0 ? L_ : a.L;
    const int M = L_ > 0 ? M_ : (L - 1) >> 1;
    const int nstrips = (nx + R - 1) / R;
    const int nxr = nstrips * R;
    // padded row: M halo + nxr + M halo + 1 slot read by the final window slide
    const int Q = (nxr + 2 * M + R) / R;
    const int rowlen = Q * R;
    const int W = it.nw;
    const int nrows = S * W;
    if (W * nstrips > C * NT)
        return;  // the planner never builds such a task
    // depth steps of this launch: [lb, le) of the task's nz (a plan may split
    // the depth range over several launches, and a round of a continuous
    // batch starts each group at its own step z0; the fields carry the state)
    const int lb = gp->z0 + a.l0;
    const int le = min(nz, gp->z0 + a.l1);
    if (lb >= le)
        return;  // this task's depth range ended in an earlier launch
    // With double_rows the task has two copies of its rows and of its image
    // partial rows: step l reads copy l % 2 and writes copy (l + 1) % 2, so
    // one barrier per step separates the reads from the next writes.
    // (two strips per thread means rows of more than 4096 positions: one copy)
    const int nbuf = (C == 1 && a.double_rows) ? 2 : 1;
    const size_t rows_size = (size_t)nrows * rowlen;  // 16-byte elements per copy
    float4 *rows = lds4;  // row (s * W + w): shot s, frequency j0 + w
    float *imgrow0 = reinterpret_cast<float *>(lds4 + nbuf * rows_size);

    // Initial rows (every copy): interior from the fields, zero halos and padding.
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
        const float4 v4 = make_float4(vf.x, vf.y, vb.x, vb.y);
        rows[row * rowlen + swz<R>(p, Q)] = v4;
        if (nbuf == 2)
            rows[rows_size + row * rowlen + swz<R>(p, Q)] = v4;
    }
    const bool shared_img = gp->img_shared != 0;
    // XP (direct atomics, compiled for the one-shot 1024-thread shapes): the
    // values of single-frequency tasks go through the wave-private transpose
    // (a wavefront's slots are consecutive strips of the one row), and
    // consecutive tasks add into different image planes, since tasks running
    // side by side would all hit the same image row at the same depth.
    const bool xpose = XP && W == 1;
    float *const imgd = (XP && a.nplanes > 1)
                            ? a.planes + (long long)(blockIdx.x % a.nplanes) * a.plane_stride
                            : a.img;
    // image sums from the rows (img_lds 2) add into a plane too when the plan has them
    float *const imgf = (ROWS_IMG && a.nplanes > 1)
                            ? a.planes + (long long)(blockIdx.x % a.nplanes) * a.plane_stride
                            : a.img;
    // image partial rows in LDS: one per (image, frequency row); every entry
    // x < nxr is written by its owner thread each step before it is read
    const int img_rows = shared_img ? 1 : nshots;
    const size_t img_size = (size_t)img_rows * W * nxr;  // floats per copy

    long long ioff[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        ioff[s] = gp->img_off[s];
    const int img_ld = gp->img_ld;
    const int slow_ld = gp->slow_ld;
    // image row 0 of the sums from the rows (read once, not per step)
    float *const imgf_rows = imgf + ioff[0];
    // a task's image sums: float atomics, or (XACC, compiled apart so the
    // float-atomic kernels are untouched) exact integer accumulation into
    // a.xacc, after which the image does not depend on the order of the adds
    auto img_add = [&](float *p, float v) {
        if constexpr (XACC)
            exact_add(a.xacc, a.xstride, (long long)(p - a.img), v);
        else
            unsafeAtomicAdd(p, v);
    };
    const float *slow0 = a.slow + gp->slow_off;
    // tables have rows of Lp = L + 1 entries (16-byte aligned): taps are
    // gathered in pairs, one 16-byte load per pair and table
    const float2 *__restrict__ TF = a.TF;
    const float2 *__restrict__ TB = a.TB;
    const int Lp = L + 1;
    // byte offset of table row q is q * Lp8; the lookup returns -q
    const int nLp8 = -8 * Lp;

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
    // slowness of the coming step at each strip's x, loaded a step ahead
    constexpr bool PREFETCH = C == 1 || CONJ;
    float sv_[C][R];
    unsigned soff[R];  // byte offsets of the first strip's x in a slowness row
#pragma unroll
    for (int r = 0; r < R; ++r)
        soff[r] = (unsigned)min(x0v[0] + r, nx - 1) * 4u;
#pragma unroll
    for (int cs = 0; cs < C; ++cs) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            sv_[cs][r] = (PREFETCH && actv[cs]) ? at_bytes(slow0 + (long long)lb * slow_ld,
                                             (unsigned)min(x0v[cs] + r, nx - 1) * 4u)
                                  : 0.f;
    }
    for (int l = lb; l < le; ++l) {
        const int cur = nbuf == 2 ? ((l - lb) & 1) : 0;
        const float4 *rows_r = rows + cur * rows_size;               // read this step
        float4 *rows_w = rows + (nbuf == 2 ? cur ^ 1 : 0) * rows_size; // written this step
        float *imgrow = imgrow0 + cur * img_size;
        // one depth step of strip cs (a compile-time index: the accumulators
        // stay in registers)
        auto strip_step = [&](auto cs_c) {
        constexpr int cs = decltype(cs_c)::value;
        // this step's imaging values for the wave-private transpose (zero on
        // lanes without a strip, which take part in the exchange)
        float xv[S][R];
        if constexpr (XP) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    xv[s][r] = 0.f;
            }
        }
        if (actv[cs]) {
            float2 (&accF)[S][R] = accF_[cs];
            float2 (&accB)[S][R] = accB_[cs];
            const int x0 = x0v[cs];
            const float om = omv[cs];
            // The slowness of this step was loaded during the previous one,
            // and the next step's load is issued now so its latency hides
            // behind the taps (at the start of a step, after the barrier,
            // every wavefront would wait for it at once). Two strips per
            // thread with separate backward tables have no registers to spare
            // (10-19 more values spill) and load at the start of the step.
            // byte offsets of the strip's table rows
            unsigned qo[R];
            if constexpr (PREFETCH) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    qo[r] = (unsigned)__mul24(kappa_index_neg(om, sv_[cs][r], a.kappa0, a.inv_dk, a.nk), nLp8);
                // One shot per group: the last step loads its own row again,
                // with no branch, so the loads go straight into the slowness
                // registers (groups of shots keep the branch: their register
                // allocation does better with it).
                if (S == 1 || l + 1 < le) {
                    const float *srow = slow0 + (long long)min(l + 1, le - 1) * slow_ld;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        // opaque to loop-invariant code motion: the offset is
                        // widened here, next to the load, which then takes the
                        // row as its scalar base (no 64-bit address per lane);
                        // the second strip forms its offsets here each step
                        unsigned o;
                        if constexpr (cs == 0) {
                            asm volatile("" : "+v"(soff[r]));
                            o = soff[r];
                        } else {
                            o = (unsigned)min(x0 + r, nx - 1) * 4u;
                            asm volatile("" : "+v"(o));
                        }
                        sv_[cs][r] = at_bytes(srow, o);
                    }
                }
            } else {
                const float *srow = slow0 + (long long)l * slow_ld;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    qo[r] = (unsigned)__mul24(
                        kappa_index_neg(om, at_bytes(srow, (unsigned)min(x0 + r, nx - 1) * 4u), a.kappa0,
                                        a.inv_dk, a.nk),
                        nLp8);
            }
            // Kappa rows are piecewise constant along x. When every active lane
            // of the wavefront has one row over its strip (mode 1), or one
            // change (mode 2: a prefix on row qo[0], the rest on qo[R-1]), the
            // strip's coefficients come from one or two gathers per tap instead
            // of R (mode 0). Every mode uses the same coefficients.
            bool sel[R];
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    accF[s][r] = make_float2(0.f, 0.f);
                    accB[s][r] = make_float2(0.f, 0.f);
                }
            }
            // x0 is a multiple of R, so p = x0 + j is at sub-row (j mod R),
            // slot x0 / R + j / R: one base per (shot, sub-row), and with a
            // fixed tap count every load is that base plus an immediate.
            const int xs = x0 / R;
            const float4 *bs[S][R];
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    bs[s][r] = rows_r + (size_t)s * shot_rows + (size_t)wv[cs] * rowlen + r * Q + xs;
            }
            auto ldw = [&](int s, int j) { return bs[s][j & (R - 1)][j >> LOG2R]; };
            auto step = [&](auto mode_c) {
                constexpr int MODE = decltype(mode_c)::value;
                // taps k, k+1 of the row at byte offset o
                auto ld4 = [](const float2 *T, unsigned o, int k) {
                    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(T) + o + 8 * k);
                };
                auto ld2 = [](const float2 *T, unsigned o, int k) {
                    return *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(T) + o + 8 * k);
                };
                // coefficients of taps k, k+1 (k even) for every x of the strip
                auto fetch2 = [&](const int k, float2 (&f0)[R], float2 (&f1)[R], float2 (&b0)[R],
                                  float2 (&b1)[R]) {
                    float4 va{}, vb{}, ua{}, ub{};
                    if constexpr (MODE != 0) {
                        va = ld4(TF, qo[0], k);
                        if constexpr (!CONJ)
                            ua = ld4(TB, qo[0], k);
                        if constexpr (MODE == 2) {
                            vb = ld4(TF, qo[R - 1], k);
                            if constexpr (!CONJ)
                                ub = ld4(TB, qo[R - 1], k);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        float4 v, u{};
                        if constexpr (MODE == 0) {
                            v = ld4(TF, qo[r], k);
                            if constexpr (!CONJ)
                                u = ld4(TB, qo[r], k);
                        } else if constexpr (MODE == 1) {
                            v = va;
                            u = ua;
                        } else {
                            // x0 is on row qo[0] and x0 + R - 1 on row qo[R - 1]
                            const bool lo = r == 0 || (r != R - 1 && sel[r]);
                            v = lo ? va : vb;
                            if constexpr (!CONJ)
                                u = lo ? ua : ub;
                        }
                        f0[r] = make_float2(v.x, v.y);
                        f1[r] = make_float2(v.z, v.w);
                        if (CONJ) {
                            b0[r] = make_float2(v.x, -v.y);
                            b1[r] = make_float2(v.z, -v.w);
                        } else {
                            b0[r] = make_float2(u.x, u.y);
                            b1[r] = make_float2(u.z, u.w);
                        }
                    }
                };
                // coefficient of tap k for every x of the strip
                auto fetch1 = [&](const int k, float2 (&f0)[R], float2 (&b0)[R]) {
                    float2 va{}, vb{}, ua{}, ub{};
                    if constexpr (MODE != 0) {
                        va = ld2(TF, qo[0], k);
                        if constexpr (!CONJ)
                            ua = ld2(TB, qo[0], k);
                        if constexpr (MODE == 2) {
                            vb = ld2(TF, qo[R - 1], k);
                            if constexpr (!CONJ)
                                ub = ld2(TB, qo[R - 1], k);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if constexpr (MODE == 0) {
                            f0[r] = ld2(TF, qo[r], k);
                            b0[r] = CONJ ? make_float2(f0[r].x, -f0[r].y) : ld2(TB, qo[r], k);
                        } else {
                            const bool lo = MODE == 1 || r == 0 || (r != R - 1 && sel[r]);
                            f0[r] = lo ? va : vb;
                            b0[r] = CONJ ? make_float2(f0[r].x, -f0[r].y) : (lo ? ua : ub);
                        }
                    }
                };
                // Wide windows, and the 128 registers of 1024-thread workgroups:
                // keep the next taps' loads from being hoisted over this one.
                // That bounds the live registers; without it the one-shot
                // 1024-thread shapes spill (24-146 values) and run 4-26 % slower.
                auto fence = [] {
                    if constexpr (R >= 8 || MAXT == 1024)
                        __builtin_amdgcn_sched_barrier(0);
                };
                if constexpr (!SYM) {
                    // new[x] = sum_k T[k] old[x + k - M]: a window of R inputs slides
                    // right by one per tap. Taps go in chunks of R so the window
                    // registers are back in place after each chunk.
                    float4 wv4[S][R];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            wv4[s][r] = ldw(s, r);
                    }
                    auto tap = [&](const int k, const float2 (&cf)[R], const float2 (&cb)[R]) {
#pragma unroll
                        for (int s = 0; s < S; ++s) {
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                accF[s][r] = mac(accF[s][r], cf[r], make_float2(wv4[s][r].x, wv4[s][r].y));
                                accB[s][r] = mac(accB[s][r], cb[r], make_float2(wv4[s][r].z, wv4[s][r].w));
                            }
                        }
#pragma unroll
                        for (int s = 0; s < S; ++s) {
#pragma unroll
                            for (int r = 0; r + 1 < R; ++r)
                                wv4[s][r] = wv4[s][r + 1];
                            wv4[s][R - 1] = ldw(s, R + k);  // the slot past the last tap is padding
                        }
                    };
                    auto tap2 = [&](const int k, const bool second) {
                        float2 f0[R], f1[R], b0[R], b1[R];
                        fetch2(k, f0, f1, b0, b1);
                        tap(k, f0, b0);
                        if (second)
                            tap(k + 1, f1, b1);
                        fence();
                    };
                    auto tap1 = [&](const int k) {
                        float2 f0[R], b0[R];
                        fetch1(k, f0, b0);
                        tap(k, f0, b0);
                        fence();
                    };
                    auto chunk = [&](const int c) {
                        if constexpr (PAIRED) {
#pragma unroll
                            for (int r2 = 0; r2 < R; r2 += 2)
                                tap2(c * R + r2, true);
                        } else {
#pragma unroll
                            for (int r2 = 0; r2 < R; ++r2)
                                tap1(c * R + r2);
                        }
                    };
                    if constexpr (L_ > 0) {
                        constexpr int NCH = L_ / R;  // R is even: chunks start at even taps
#pragma unroll
                        for (int c = 0; c < NCH; ++c)
                            chunk(c);
                        if constexpr (PAIRED) {
#pragma unroll
                            for (int k = NCH * R; k < L_; k += 2)
                                tap2(k, k + 1 < L_);
                        } else {
#pragma unroll
                            for (int k = NCH * R; k < L_; ++k)
                                tap1(k);
                        }
                    } else {
                        const int nch = L / R;
#pragma unroll 1
                        for (int c = 0; c < nch; ++c)
                            chunk(c);
                        if constexpr (PAIRED) {
#pragma unroll 1
                            for (int k = nch * R; k < L; k += 2)
                                tap2(k, k + 1 < L);
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
                    float4 lw[S][R], rw[S][R];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            lw[s][r] = ldw(s, r);
                            rw[s][r] = ldw(s, r + 2 * M);
                        }
                    }
                    auto pair = [&](const int k, const float2 (&cf)[R], const float2 (&cb)[R]) {
#pragma unroll
                        for (int s = 0; s < S; ++s) {
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                const float4 u = lw[s][r], v = rw[s][r];
                                accF[s][r] = cmac(accF[s][r], cf[r], make_float2(u.x + v.x, u.y + v.y));
                                accB[s][r] = cmac(accB[s][r], cb[r], make_float2(u.z + v.z, u.w + v.w));
                            }
                        }
#pragma unroll
                        for (int s = 0; s < S; ++s) {
#pragma unroll
                            for (int r = 0; r + 1 < R; ++r)
                                lw[s][r] = lw[s][r + 1];
                            lw[s][R - 1] = ldw(s, R + k);
#pragma unroll
                            for (int r = R - 1; r > 0; --r)
                                rw[s][r] = rw[s][r - 1];
                            rw[s][0] = ldw(s, 2 * M - 1 - k);
                        }
                    };
                    auto pair2 = [&](const int k, const bool second) {
                        float2 f0[R], f1[R], b0[R], b1[R];
                        fetch2(k, f0, f1, b0, b1);
                        pair(k, f0, b0);
                        if (second)
                            pair(k + 1, f1, b1);
                        fence();
                    };
                    auto pair1 = [&](const int k) {
                        float2 f0[R], b0[R];
                        fetch1(k, f0, b0);
                        pair(k, f0, b0);
                        fence();
                    };
                    auto chunk = [&](const int c) {
                        if constexpr (PAIRED) {
#pragma unroll
                            for (int r2 = 0; r2 < R; r2 += 2)
                                pair2(c * R + r2, true);
                        } else {
#pragma unroll
                            for (int r2 = 0; r2 < R; ++r2)
                                pair1(c * R + r2);
                        }
                    };
                    if constexpr (L_ > 0) {
                        constexpr int NCH = M_ / R;
#pragma unroll
                        for (int c = 0; c < NCH; ++c)
                            chunk(c);
                        if constexpr (PAIRED) {
#pragma unroll
                            for (int k = NCH * R; k < M_; k += 2)
                                pair2(k, k + 1 < M_);
                        } else {
#pragma unroll
                            for (int k = NCH * R; k < M_; ++k)
                                pair1(k);
                        }
                    } else {
                        const int nch = M / R;
#pragma unroll 1
                        for (int c = 0; c < nch; ++c)
                            chunk(c);
                        if constexpr (PAIRED) {
#pragma unroll 1
                            for (int k = nch * R; k < M; k += 2)
                                pair2(k, k + 1 < M);
                        } else {
#pragma unroll 1
                            for (int k = nch * R; k < M; ++k)
                                pair1(k);
                        }
                    }
                    float2 cf[R], cb[R];
                    fetch1(M, cf, cb);
#pragma unroll
                    for (int s = 0; s < S; ++s) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            accF[s][r] = cmac(accF[s][r], cf[r], make_float2(lw[s][r].x, lw[s][r].y));
                            accB[s][r] = cmac(accB[s][r], cb[r], make_float2(lw[s][r].z, lw[s][r].w));
                        }
                    }
                }
            };
            // The mode is uniform over the wavefront (a scalar branch). Groups of
            // several shots already share each gather over the shots and are
            // VALU-bound: they keep one gather per x and its register budget.
            if constexpr (S == 1) {
                // mode 1: one row over the strip (lowest row == highest row)
                unsigned qlo = qo[0], qhi = qo[0];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    qlo = min(qlo, qo[r]);
                    qhi = max(qhi, qo[r]);
                }
                if (a.strip_gathers && __ballot(qlo != qhi) == 0) {
                    step(std::integral_constant<int, 1>{});
                } else {
                    int changes = 0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        sel[r] = qo[r] == qo[0];
                        if (r > 0)
                            changes += qo[r] != qo[r - 1] ? 1 : 0;
                    }
                    if (a.strip_gathers && __ballot(changes > 1) == 0)
                        step(std::integral_constant<int, 2>{});
                    else
                        step(std::integral_constant<int, 0>{});
                }
            } else {
                step(std::integral_constant<int, 0>{});
            }
            // imaging condition of this step. With img_lds 1 the thread stores
            // its R values (summed over the shots when they share an image)
            // in the partial row of its frequency; the flush below sums the
            // task's frequency rows in fixed order, one atomic per (depth, x).
            // With img_lds 2 the flush forms the terms itself from the new
            // rows (nothing to do here).
            if (ROWS_IMG && a.img_lds == 2) {
            } else if (a.img_lds) {
                const int w = wv[cs];
                if (shared_img) {
                    float v[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        v[r] = rmc(accF[0][r], accB[0][r]);  // a group has at least one shot
#pragma unroll
                        for (int s = 1; s < S; ++s)
                            if (s < nshots)
                                v[r] += rmc(accF[s][r], accB[s][r]);
                    }
                    store_row<R>(imgrow + (size_t)w * nxr + x0, v);
                } else {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        if (s < nshots) {
                            float v[R];
#pragma unroll
                            for (int r = 0; r < R; ++r)
                                v[r] = rmc(accF[s][r], accB[s][r]);
                            store_row<R>(imgrow + ((size_t)s * W + w) * nxr + x0, v);
                        }
                    }
                }
            } else if (xpose) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (s < nshots)
                            xv[s][r] = rmc(accF[s][r], accB[s][r]);
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s < nshots) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int x = x0 + r;
                            if (x < nx)
                                unsafeAtomicAdd(imgd + ioff[s] + (long long)l * img_ld + x,
                                                rmc(accF[s][r], accB[s][r]));
                        }
                    }
                }
            }
        }
        if constexpr (XP) {
        if (xpose) {
            // Direct atomics of a single-frequency task: the wavefront's 64
            // strips are 64*R consecutive x from x_w. Lane i stores its R values
            // at seg[i*R ...], then lane j adds seg[j + 64k] to x = x_w + j + 64k
            // (k < R): each atomic instruction covers 64 consecutive floats.
            // The bound is re-read each step (the empty asm keeps the compiler
            // from hoisting the R compares out of the depth loop, where their
            // masks would occupy registers through the taps).
            const int lane = tid & 63;
            float *seg = imgrow0 + (tid >> 6) * (64 * R);
            const int xl = ((tid & ~63) + cs * NT) * R + lane;  // x of k = 0
            int nxl = nx;
            asm volatile("" : "+v"(nxl));
            const long long roff = (long long)l * img_ld;
            const int nimg = shared_img ? 1 : nshots;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (s < nimg) {
                    float v[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if (shared_img) {
                            v[r] = xv[0][r];
#pragma unroll
                            for (int t = 1; t < S; ++t)
                                v[r] += xv[t][r];
                        } else {
                            v[r] = xv[s][r];
                        }
                    }
                    wave_sync();  // the previous reads of seg are done
                    store_row<R>(seg + lane * R, v);
                    wave_sync();
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        const float u = seg[lane + 64 * k];
                        if (xl + 64 * k < nxl)
                            unsafeAtomicAdd(imgd + ioff[s] + roff + xl + 64 * k, u);
                    }
                }
            }
        }
        }
        };
        // position x0 + r + M of a row is at sub-row (r + M) mod R, slot
        // x0 / R + (r + M) / R (x0 is a multiple of R): the per-r part is the
        // same for every thread
        auto store_strip = [&](auto cs_c) {
            constexpr int cs = decltype(cs_c)::value;
            const int sbase = wv[cs] * rowlen + (x0v[cs] >> LOG2R);
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int x = x0v[cs] + r;
                    if (x < nx) {
                        const float2 f = accF_[cs][s][r], b = accB_[cs][s][r];
                        const int c = ((r + M) & (R - 1)) * Q + ((r + M) >> LOG2R);
                        rows_w[s * shot_rows + sbase + c] = make_float4(f.x, f.y, b.x, b.y);
                    }
                }
            }
        };
        // img_lds 2 (one shot per group): Re(F conj B) of the rows written
        // this step, summed per x over the task's frequency rows in row order,
        // one atomic per (depth, x). Thread t takes x = t, t + NT, ...; the
        // slot of x + NT is that of x plus NT / R (NT is a multiple of R).
        auto flush_rows = [&](const float4 *src, int ld_l) {
            float *dst = imgf_rows + (long long)ld_l * img_ld;
            // the slot of x = tid, formed each step from an opaque copy of tid:
            // a few integer operations instead of a value kept through the taps,
            // where it would be spilled and reloaded with a memory wait per step
            // (the slot is opaque too, so the rows' offsets stay additions)
            int t = tid;
            asm volatile("" : "+v"(t));
            int slot = __mul24((t + M) & (R - 1), Q) + ((t + M) >> LOG2R);
            asm volatile("" : "+v"(slot));
            const float4 *e0 = src + slot;
            const unsigned xend = (unsigned)nx * 4u;
            // the first row's term, and the sum so far plus a further row's
            auto first = [&](const float4 g) {
                if constexpr (EXACT)
                    return rmc(make_float2(g.x, g.y), make_float2(g.z, g.w));
                else
                    return fmaf(g.y, g.w, g.x * g.z);
            };
            auto more = [&](float v, const float4 g) {
                if constexpr (EXACT)
                    return v + rmc(make_float2(g.x, g.y), make_float2(g.z, g.w));
                else
                    return fmaf(g.y, g.w, fmaf(g.x, g.z, v));
            };
            if (W == 1) {
                // single-frequency tasks (the wide apertures): one read per x
                for (unsigned xo = (unsigned)t * 4u; xo < xend; xo += 4u * NT, e0 += NT >> LOG2R)
                    img_add(reinterpret_cast<float *>(reinterpret_cast<char *>(dst) + xo), first(*e0));
                return;
            }
            for (unsigned xo = (unsigned)t * 4u; xo < xend; xo += 4u * NT, e0 += NT >> LOG2R) {
                const float4 *e = e0;
                float v = first(*e);
                // rows in order; four reads issued before their sums, so the
                // LDS latency is paid once per four rows
                int w = 1;
                for (; w + 4 <= W; w += 4, e += 4 * rowlen) {
                    const float4 g0 = e[rowlen], g1 = e[2 * rowlen], g2 = e[3 * rowlen],
                                 g3 = e[4 * rowlen];
                    v = more(more(more(more(v, g0), g1), g2), g3);
                }
                for (; w < W; ++w) {
                    e += rowlen;
                    v = more(v, *e);
                }
                img_add(reinterpret_cast<float *>(reinterpret_cast<char *>(dst) + xo), v);
            }
        };
        using c0 = std::integral_constant<int, 0>;
        using c1 = std::integral_constant<int, 1>;
        strip_step(c0{});
        if (nbuf == 2 && actv[0])
            store_strip(c0{});
        if constexpr (C == 2) {
            strip_step(c1{});
            if (nbuf == 2 && actv[1])
                store_strip(c1{});
        }
        static_assert(C == 1 || C == 2, "one or two strips per thread");
        __syncthreads();
        if (nbuf == 1) {
            if (actv[0])
                store_strip(c0{});
            if constexpr (C == 2)
                if (actv[1])
                    store_strip(c1{});
        }
        // With two copies the rows written this step are complete after the
        // barrier and stay unchanged until the barrier of the next step.
        if constexpr (ROWS_IMG)
            if (a.img_lds == 2 && nbuf == 2)
                flush_rows(rows_w, l);
        if (a.img_lds == 1) {
            // consecutive lanes flush consecutive x: one 32-byte sector of the
            // image carries 8 atomic adds
            for (int s = 0; s < img_rows; ++s) {
                float *dst = a.img + ioff[s] + (long long)l * img_ld;
                const float *src = imgrow + (size_t)s * W * nxr;
                for (int x = tid; x < nx; x += NT) {
                    float v = src[x];
                    for (int w = 1; w < W; ++w)
                        v += src[(size_t)w * nxr + x];
                    img_add(dst + x, v);
                }
            }
        }
        if (nbuf == 1) {
            __syncthreads();
            // one copy: the new rows are complete after the second barrier;
            // the next step writes them only after its first barrier
            if constexpr (ROWS_IMG)
                if (a.img_lds == 2)
                    flush_rows(rows, l);
        }
    }

    if (a.write_back) {
        // after the last step's barrier; with two copies the last step wrote
        // copy (le - lb) % 2
        const float4 *fin = rows + (nbuf == 2 ? ((le - lb) & 1) : 0) * rows_size;
        for (int e = tid; e < nrows * nx; e += NT) {
            const int row = e / nx;
            const int x = e - row * nx;
            const int s = row / W;
            const int ww = row - s * W;
            if (s < nshots) {
                const float4 v = fin[row * rowlen + swz<R>(x + M, Q)];
                const long long r = (long long)(it.j0 + ww) * ld + x;
                a.F[gp->f_off[s] + r] = make_float2(v.x, v.y);
                a.B[gp->b_off[s] + r] = make_float2(v.z, v.w);
            }
        }
    }
}

template <int L_, int R, int S, bool CONJ, bool SYM, bool PAIRED, int C, int MAXT, bool XP,
          bool EXACT, bool XACC>
hipError_t launch_fused_k(int n_items, int block, size_t lds, hipStream_t st, const FusedArgs &a)
{
    auto k = fused_kernel<L_, R, S, CONJ, SYM, PAIRED, C, MAXT, XP, EXACT, XACC>;
    hipError_t e = ensure_lds(k, lds);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k, dim3(n_items), dim3(block), lds, st, a);
    return hipGetLastError();
}

// exact image sums (a.xacc) are compiled for the shapes that flush task sums
// (no wave-private transpose) with conjugate tables
template <int L_, int R, int S, bool CONJ, bool SYM, bool PAIRED, int C, int MAXT, bool XP,
          bool EXACT>
hipError_t launch_fused_t(int n_items, int block, size_t lds, hipStream_t st, const FusedArgs &a)
{
    if constexpr (!XP && CONJ) {
        if (a.xacc)
            return launch_fused_k<L_, R, S, CONJ, SYM, PAIRED, C, MAXT, XP, EXACT, true>(n_items, block,
                                                                                       lds, st, a);
    } else {
        if (a.xacc)
            return hipErrorInvalidValue;
    }
    return launch_fused_k<L_, R, S, CONJ, SYM, PAIRED, C, MAXT, XP, EXACT, false>(n_items, block, lds,
                                                                                  st, a);
}

template <int L_, bool P, int R, int S, int C, int MAXT, bool XP = false>
hipError_t launch_fused_RS(const FusedShape &sh, int n_items, size_t lds, hipStream_t st,
                           const FusedArgs &a)
{
    const int b = sh.block;
    if (sh.sym)
        return sh.conj ? launch_fused_t<L_, R, S, true, true, P, C, MAXT, XP, false>(n_items, b, lds, st, a)
                       : launch_fused_t<L_, R, S, false, true, P, C, MAXT, XP, false>(n_items, b, lds, st, a);
    return sh.conj ? launch_fused_t<L_, R, S, true, false, P, C, MAXT, XP, false>(n_items, b, lds, st, a)
                   : launch_fused_t<L_, R, S, false, false, P, C, MAXT, XP, false>(n_items, b, lds, st, a);
}

// The CPU path's rounding: taps in order (no pairing), compiled for the
// one-shot 1024-thread shapes only.
template <int L_, bool P, int C, bool XP>
hipError_t launch_fused_exact(const FusedShape &sh, int n_items, size_t lds, hipStream_t st,
                              const FusedArgs &a)
{
    const int b = sh.block;
    return sh.conj ? launch_fused_t<L_, 4, 1, true, false, P, C, 1024, XP, true>(n_items, b, lds, st, a)
                   : launch_fused_t<L_, 4, 1, false, false, P, C, 1024, XP, true>(n_items, b, lds, st, a);
}

template <int L_, bool P>
hipError_t launch_fused_L(const FusedShape &sh, int n_items, size_t lds, hipStream_t st,
                          const FusedArgs &a)
{
    if (sh.exact) {
        if (!(sh.R == 4 && sh.S == 1 && sh.block > 512))
            return hipErrorInvalidValue;
        const bool xp = a.img_lds == 0 && a.img_xpose;
        if (sh.C == 1)
            return xp ? launch_fused_exact<L_, P, 1, true>(sh, n_items, lds, st, a)
                      : launch_fused_exact<L_, P, 1, false>(sh, n_items, lds, st, a);
        if (sh.C == 2)
            return xp ? launch_fused_exact<L_, P, 2, true>(sh, n_items, lds, st, a)
                      : launch_fused_exact<L_, P, 2, false>(sh, n_items, lds, st, a);
        return hipErrorInvalidValue;
    }
    if (sh.R == 4 && sh.S == 1 && sh.block > 512) {  // 16 waves: registers for 4 per SIMD
        const bool xp = a.img_lds == 0 && a.img_xpose;
        if (sh.C == 1)
            return xp ? launch_fused_RS<L_, P, 4, 1, 1, 1024, true>(sh, n_items, lds, st, a)
                      : launch_fused_RS<L_, P, 4, 1, 1, 1024>(sh, n_items, lds, st, a);
        if (sh.C == 2)
            return xp ? launch_fused_RS<L_, P, 4, 1, 2, 1024, true>(sh, n_items, lds, st, a)
                      : launch_fused_RS<L_, P, 4, 1, 2, 1024>(sh, n_items, lds, st, a);
        return hipErrorInvalidValue;
    }
    if (sh.R == 2 && sh.S == 2 && sh.C == 1 && sh.block > 512)
        return launch_fused_RS<L_, P, 2, 2, 1, 1024>(sh, n_items, lds, st, a);
    if (sh.block > 512)
        return hipErrorInvalidValue;
    if (sh.C == 1) {
        if (sh.R == 4 && sh.S == 1)
            return launch_fused_RS<L_, P, 4, 1, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 4 && sh.S == 2)
            return launch_fused_RS<L_, P, 4, 2, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 4 && sh.S == 4)
            return launch_fused_RS<L_, P, 4, 4, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 8 && sh.S == 1)
            return launch_fused_RS<L_, P, 8, 1, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 2 && sh.S == 2)
            return launch_fused_RS<L_, P, 2, 2, 1, 512>(sh, n_items, lds, st, a);
        if (sh.R == 2 && sh.S == 4)
            return launch_fused_RS<L_, P, 2, 4, 1, 512>(sh, n_items, lds, st, a);
    } else if (sh.C == 2) {
        if (sh.R == 8 && sh.S == 1)
            return launch_fused_RS<L_, P, 8, 1, 2, 512>(sh, n_items, lds, st, a);
    }
    return hipErrorInvalidValue;
}

}  // namespace

#define WLSQ_FUSED_NAME2(l, p) launch_fused_L##l##_P##p
#define WLSQ_FUSED_NAME(l, p) WLSQ_FUSED_NAME2(l, p)
hipError_t WLSQ_FUSED_NAME(WLSQ_L, WLSQ_PAIRED)(const FusedShape &sh, int n_items, size_t lds,
                                                hipStream_t st, const FusedArgs &a)
{
    return launch_fused_L<WLSQ_L, WLSQ_PAIRED != 0>(sh, n_items, lds, st, a);
}

}  // namespace wlsq
