// Internal declarations shared by the kernels and the C ABI of libwlsq_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "wlsq_hip.h"

namespace wlsq {

constexpr int kMaxShotsPerGroup = 4;
// Largest kappa sample count: below 2^22 the rounding step of the lookup is
// exact (u - 0.5 and 0.5 - u are representable for 0 <= u < 2^22).
constexpr int kMaxKappaSamples = 1 << 22;
constexpr int kMaxImagePlanes = 16;
// Most depth steps of one round of a continuous batch (wlsq_batch_round_ws).
constexpr int kMaxRoundSteps = 1 << 24;
// Exact image accumulation: 64-bit words per image point (16 digit words of
// 16-bit spacing from 2^-149, then a flags word).
constexpr int kExactWords = 16;
constexpr int kExactPlanes = kExactWords + 1;

// One shot group: up to 4 shots that share aperture, depth count and
// frequency list, so one coefficient load serves all of them.
struct Group {
    long long f_off[kMaxShotsPerGroup];   // row 0 of each shot's forward field
    long long b_off[kMaxShotsPerGroup];   // row 0 of each shot's backward field
    long long img_off[kMaxShotsPerGroup]; // image row 0 of each shot
    long long slow_off;                   // slowness (depth 0, window column 0)
    int nshots;                           // 1..S
    int ld;                               // field row stride
    int img_ld;                           // image row stride
    int slow_ld;                          // slowness row stride
    int nx;                               // aperture
    int nz;                               // depth steps
    int nf;                               // frequencies
    int omega_off;                        // first frequency in the omega pool
    int img_shared;                       // all shots add into the same image rows
    int z0;                               // first depth step: a launch runs [z0 + l0, z0 + l1)
};

// One workgroup task: `nw` consecutive frequencies (from `j0`) of one group,
// the depth steps of the launch.
struct Item {
    int group;
    int j0;
    int nw;
    int cost_rank;
};

struct FusedArgs {
    const Item *items;
    const Group *groups;
    float2 *F;
    float2 *B;
    float *img;
    const float *omega;
    const float *slow;
    const float2 *TF;  // row-padded copies (nk, L + 1) of the compact tables
    const float2 *TB;
    int nk;
    float kappa0;
    float inv_dk;
    int L;
    int write_back;
    int img_lds;  // 1: reduce image rows in LDS, flush with lane-contiguous atomics
    int img_xpose;  // img_lds 0: single-frequency tasks transpose each wavefront's values
                    // through a wave-private LDS segment (lane-contiguous atomics)
    float *planes;  // img_lds 0 and nplanes > 1: task b adds its direct atomics into
                    // plane b % nplanes (laid out like img), summed into img afterwards
    long long plane_stride;
    int nplanes;
    int strip_gathers;  // 1: one or two gathers per strip where the kappa rows allow it
    int double_rows;    // 1: two copies of the rows in LDS, one barrier per step
    int l0, l1;         // depth steps [z0 + l0, min(nz, z0 + l1)) of each task in this
                        // launch (z0 of the task's group)
    unsigned long long *xacc;  // non-null: the image sums go into these exact accumulators
    long long xstride;         // (kExactPlanes planes laid out like img, `xstride` apart)
};

// Launch shape of the fused kernel; chosen by the planner.
struct FusedShape {
    int R;      // consecutive x per thread (2, 4, 8)
    int S;      // shots per group (1, 2, 4)
    int C;      // strips per thread (1, or 2: R = 8, or R = 4 above 512 threads)
    int block;  // threads per workgroup
    bool conj;  // table_back == conj(table_forw): load one table
    bool sym;   // tables are tap-symmetric (T[k] == T[L-1-k]): pair the taps
    bool paired;  // gather two taps per 16-byte load (fewer address-unit cycles)
    bool exact;   // the CPU path's rounding: separate multiplies and adds, taps in order
};

// LDS bytes one task needs: two fields x S shots x W rows, swizzled rows of
// R sub-rows, plus `img_rows` x W image partial rows of the padded aperture;
// `copies` = 2 doubles both (one barrier per depth step).
size_t fused_lds_bytes(int nx, int M, int R, int S, int W, int img_rows, int copies = 1);

hipError_t launch_fused(const FusedShape &sh, int n_items, size_t lds_bytes, hipStream_t st,
                        const FusedArgs &a);
// One object of wlsq_fused.hip per taps (25, 31, 51, 0 = any) and gather width
// (P1: two taps per 16-byte load).
#define WLSQ_DECLARE_FUSED(l, p)                                                              \
    hipError_t launch_fused_L##l##_P##p(const FusedShape &sh, int n_items, size_t lds,        \
                                        hipStream_t st, const FusedArgs &a);
WLSQ_DECLARE_FUSED(25, 0)
WLSQ_DECLARE_FUSED(25, 1)
WLSQ_DECLARE_FUSED(31, 0)
WLSQ_DECLARE_FUSED(31, 1)
WLSQ_DECLARE_FUSED(51, 0)
WLSQ_DECLARE_FUSED(51, 1)
WLSQ_DECLARE_FUSED(0, 0)
WLSQ_DECLARE_FUSED(0, 1)
#undef WLSQ_DECLARE_FUSED

// Plain path kernels.
hipError_t launch_transpose_tapmajor(const float2 *nat, float2 *tap, long long n_slices, int nx,
                                     int L, hipStream_t st);
// Frequency groups of one plain step (blocks along omega); part holds
// [shot][depth][group][x] floats.
int plain_groups(int nf);
hipError_t launch_plain_step(int S, const float2 *opF, const float2 *opB, const float2 *oldF,
                             const float2 *oldB, long long in_shot_stride, float2 *newF,
                             float2 *newB, float *part, long long part_stride, int nshots, int nf,
                             int nx, int M, hipStream_t st);
hipError_t launch_plain_reduce(const float *part, float *image, int nshots, long long image_shot_stride,
                               int nextrap, int groups, int nx, hipStream_t st);
hipError_t launch_copy_rows(const float2 *src, long long src_shot_stride, float2 *dst,
                            long long dst_shot_stride, int nshots, int nrows, int nx,
                            hipStream_t st);
hipError_t launch_expand(const float *slow, int slow_ld, const float *omega, float kappa0,
                         float inv_dk, int nk, const float2 *table, float2 *out, int nextrap,
                         int nf, int nx, int L, int layout, hipStream_t st);
hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t st);
// blocks x 256 threads, 8 * iters FP32 FMAs per wave (counter calibration)
hipError_t launch_probe_fma(int blocks, int iters, float *out, hipStream_t st);
// `bytes` (a multiple of 4) of host memory to `dst` through kernel arguments
// (chunks under 4 KB): no host-to-device copy
hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t st);
// img[i] += sum over c < nplanes of planes[c * stride + i] (fixed order), for
// i < n where the sum is nonzero
hipError_t launch_planes_reduce(const float *planes, long long stride, int nplanes, float *img,
                                long long n, hipStream_t st);
// compact (nk, L) table -> rows of L + 1 entries (last 0), read by the fused kernel
hipError_t launch_table_pad(const float2 *in, float2 *out, int nk, int L, hipStream_t st);
// img[i] += the float nearest the exact sum in acc (words `stride` apart), for
// i < n where any term was added
hipError_t launch_exact_finish(const unsigned long long *acc, long long stride, long long n,
                               float *img, hipStream_t st);
// out[j] = the float nearest sum_i terms[i * m + j] (i < n), the terms added by
// many threads in no fixed order into zeroed accumulators acc (kExactPlanes * m words)
hipError_t launch_exact_probe(const float *terms, long long n, int m, unsigned long long *acc,
                              float *out, hipStream_t st);

// Records the message returned by wlsq_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);

// (float)(1.0 / (double)dkappa): the reciprocal every implementation of the
// kappa lookup uses.
inline float inv_dkappa(float dkappa) { return (float)(1.0 / (double)dkappa); }

}  // namespace wlsq
