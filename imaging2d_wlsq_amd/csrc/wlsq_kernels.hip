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
//  * fused (k-table): wlsq_fused.hip; this file holds its dispatch by taps
//    and gather width.
#include "wlsq_device.h"

#include <algorithm>
#include <cstring>

namespace wlsq {
namespace {

// ------------------------------------------------------------------ plain path
constexpr int kPlainTX = 64;  // x per block
constexpr int kPlainTJ = 4;   // frequencies per block

// A coefficient of the expanded tables: read once per launch, so the load is
// marked nontemporal (it does not displace the wavefield tiles from the caches;
// 3-5 % more table bandwidth on MI355X).
__device__ __forceinline__ float2 load_stream(const float2 *p)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(p));
    return make_float2(v.x, v.y);
}

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
        auto tap = [&](const int k) {
            const float2 cf = load_stream(cF + (long long)k * nx);
            const float2 cb = load_stream(cB + (long long)k * nx);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int o = (s * kPlainTJ + ty) * TW + tx + k;
                aF[s] = cmac(aF[s], cf, smF[o]);
                aB[s] = cmac(aB[s], cb, smB[o]);
            }
        };
        // with a fixed tap count every table load of the step is issued up
        // front (8 % more table bandwidth than 8 taps at a time)
        if constexpr (L_ > 0) {
#pragma unroll
            for (int k = 0; k < L_; ++k)
                tap(k);
        } else {
#pragma unroll 8
            for (int k = 0; k < L; ++k)
                tap(k);
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

// Host bytes carried to the device in a kernel argument (no host-to-device
// copy, so nothing is allocated, staged or captured as a copy node).
constexpr int kUploadWords = 992;  // 3968 bytes: the argument stays under 4 KB
struct UploadChunk {
    unsigned nwords;
    unsigned w[kUploadWords];
};

__global__ void upload_kernel(unsigned *__restrict__ dst, const UploadChunk c)
{
    for (unsigned e = threadIdx.x; e < c.nwords; e += blockDim.x)
        dst[e] = c.w[e];
}

// Counter calibration: each thread runs `iters` rounds of eight independent
// FP32 fused multiply-add chains, so every wave issues 8 * iters FMA
// instructions in its loop, plus a short fixed tail (seven adds and the store).
__global__ void __launch_bounds__(256) probe_fma_kernel(float *out, float m, float c, int iters)
{
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        a[k] = (float)threadIdx.x * 1e-3f + (float)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            a[k] = fmaf(a[k], m, c);
    }
    out[blockIdx.x * (long long)blockDim.x + threadIdx.x] =
        ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

__global__ void fill_f32_kernel(float *p, long long n, float v)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride)
        p[e] = v;
}

// Image planes of the direct atomics summed into the image, plane 0 first.
// Positions no task touched sum to zero and keep their bits.
__global__ void planes_reduce_kernel(const float *__restrict__ planes, long long pstride,
                                     int nplanes, float *img, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride) {
        float s = 0.f;
        for (int c = 0; c < nplanes; ++c)
            s += planes[c * pstride + e];
        if (s != 0.f)
            img[e] += s;
    }
}

}  // namespace

size_t fused_lds_bytes(int nx, int M, int R, int S, int W, int img_rows, int copies)
{
    const size_t nstrips = (size_t)(nx + R - 1) / R;
    const size_t nxr = nstrips * R;
    const size_t Q = (nxr + 2 * (size_t)M + R) / R;
    const size_t rowlen = Q * R;
    return (size_t)copies * (2 * (size_t)S * W * rowlen * sizeof(float2) +
                             (size_t)img_rows * W * nxr * sizeof(float));
}

hipError_t launch_fused(const FusedShape &sh, int n_items, size_t lds_bytes, hipStream_t st,
                        const FusedArgs &a)
{
    if (n_items <= 0)
        return hipSuccess;
    if (sh.block <= 0 || sh.block > 1024 || sh.block % 64 != 0)
        return hipErrorInvalidValue;
#define WLSQ_FUSED_BY_P(l)                                                                  \
    return sh.paired ? launch_fused_L##l##_P1(sh, n_items, lds_bytes, st, a)                \
                     : launch_fused_L##l##_P0(sh, n_items, lds_bytes, st, a);
    switch (a.L) {
    case 25: WLSQ_FUSED_BY_P(25)
    case 31: WLSQ_FUSED_BY_P(31)
    case 51: WLSQ_FUSED_BY_P(51)
    default: WLSQ_FUSED_BY_P(0)
    }
#undef WLSQ_FUSED_BY_P
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

hipError_t launch_planes_reduce(const float *planes, long long stride, int nplanes, float *img,
                                long long n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    hipLaunchKernelGGL(planes_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, planes,
                       stride, nplanes, img, n);
    return hipGetLastError();
}

hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    if (bytes % 4 != 0 || reinterpret_cast<uintptr_t>(dst) % 4 != 0)
        return hipErrorInvalidValue;
    const unsigned *w = static_cast<const unsigned *>(src);
    unsigned *d = static_cast<unsigned *>(dst);
    size_t n = bytes / 4;
    for (size_t o = 0; o < n; o += kUploadWords) {
        UploadChunk c;
        c.nwords = (unsigned)std::min<size_t>(kUploadWords, n - o);
        std::memcpy(c.w, w + o, sizeof(unsigned) * c.nwords);
        hipLaunchKernelGGL(upload_kernel, dim3(1), dim3(256), 0, st, d + o, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_probe_fma(int blocks, int iters, float *out, hipStream_t st)
{
    if (blocks <= 0 || iters < 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(probe_fma_kernel, dim3(blocks), dim3(256), 0, st, out, 0.999f, 1e-3f, iters);
    return hipGetLastError();
}

// Exact image sums: img[i] += the float nearest the sum accumulated for point i
// (points no term reached keep their value).
__global__ void __launch_bounds__(256) exact_finish_kernel(const unsigned long long *__restrict__ acc,
                                                           long long stride, long long n,
                                                           float *__restrict__ img)
{
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        unsigned long long w[kExactWords];
        unsigned long long any = 0;
#pragma unroll
        for (int k = 0; k < kExactWords; ++k) {
            w[k] = acc[k * stride + i];
            any |= w[k];
        }
        const unsigned long long f = acc[kExactWords * stride + i];
        if (any | f)
            img[i] += exact_round(w, f);
    }
}

// Measurement aid: thread t of block b adds terms[i * m + j] for the i of its
// stride, so the terms of one sum arrive in no fixed order; then one thread
// per sum rounds it.
__global__ void __launch_bounds__(256) exact_probe_add_kernel(const float *__restrict__ terms,
                                                              long long n, int m,
                                                              unsigned long long *acc)
{
    const long long total = n * m;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (long long)gridDim.x * 256)
        exact_add(acc, m, e % m, terms[e]);
}

__global__ void __launch_bounds__(256) exact_probe_round_kernel(const unsigned long long *__restrict__ acc,
                                                                int m, float *__restrict__ out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m)
        return;
    unsigned long long w[kExactWords];
#pragma unroll
    for (int k = 0; k < kExactWords; ++k)
        w[k] = acc[(long long)k * m + j];
    out[j] = exact_round(w, acc[(long long)kExactWords * m + j]);
}

hipError_t launch_exact_finish(const unsigned long long *acc, long long stride, long long n,
                               float *img, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(exact_finish_kernel, dim3((unsigned)blocks), dim3(256), 0, st, acc, stride, n, img);
    return hipGetLastError();
}

hipError_t launch_exact_probe(const float *terms, long long n, int m, unsigned long long *acc,
                              float *out, hipStream_t st)
{
    if (m <= 0)
        return hipSuccess;
    hipError_t e = launch_fill_f32(reinterpret_cast<float *>(acc), 2ll * kExactPlanes * m, 0.f, st);
    if (e != hipSuccess)
        return e;
    if (n > 0) {
        long long blocks = (n * m + 255) / 256;
        if (blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(exact_probe_add_kernel, dim3((unsigned)blocks), dim3(256), 0, st, terms, n, m, acc);
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(exact_probe_round_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, acc, m, out);
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
