// C ABI of libwlsq_hip.so: argument checks, device scratch, the work-list
// planner of the fused kernel, plans and graphs. Declarations and the
// reference interfaces they replace are in include/wlsq_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "wlsq_internal.h"

using namespace wlsq;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? WLSQ_ENOMEM : WLSQ_EHIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define WLSQ_TRY(expr)      \
    do {                    \
        int r_ = (expr);    \
        if (r_ != WLSQ_OK)  \
            return r_;      \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int reserve(size_t need)
    {
        if (need <= bytes)
            return WLSQ_OK;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess)
                return fail(WLSQ_EHIP, "hipFree: %s", hipGetErrorString(e));
        }
        hipError_t e = hipMalloc(&p, need);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(WLSQ_ENOMEM, "hipMalloc(%zu bytes): %s", need, hipGetErrorString(e));
        }
        bytes = need;
        return WLSQ_OK;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Scratch of one stream: calls queued on one stream run in order, so they can
// share it; calls on different streams get different scratch.
struct Scratch {
    DevBuf tapF, tapB, pingF[2], pingB[2], part, stage, fields_f, fields_b, img, slow, omega, tf,
        tb;
};

struct PlanImpl;

struct Device {
    int ordinal = 0;
    hipDeviceProp_t props{};
    hipStream_t own_stream = nullptr;  // default priority, used by the host entries (created
                                       // by their first call)
    void *scratch_ptr = nullptr;       // this device's copy of `scratch` (allocation-free entry)
    std::mutex mu;           // scratch map, plan cache, device-pointer entries
    std::mutex host_call_mu; // one host-array entry call at a time (shared scratch)
    std::map<hipStream_t, std::unique_ptr<Scratch>> scratch;
    std::map<std::vector<long long>, std::shared_ptr<PlanImpl>> ktable_plans;
};

std::mutex g_devices_mu;
std::map<int, std::unique_ptr<Device>> g_devices;

int current_device(Device **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_devices_mu);
    auto it = g_devices.find(dev);
    if (it == g_devices.end()) {
        auto d = std::make_unique<Device>();
        d->ordinal = dev;
        HIP_TRY(hipGetDeviceProperties(&d->props, dev));
        it = g_devices.emplace(dev, std::move(d)).first;
    }
    *out = it->second.get();
    return WLSQ_OK;
}

// The stream of the host-array entries (default priority).
int host_stream(Device *d, hipStream_t *out)
{
    std::lock_guard<std::mutex> lk(d->mu);
    if (!d->own_stream)
        HIP_TRY(hipStreamCreateWithFlags(&d->own_stream, hipStreamNonBlocking));
    *out = d->own_stream;
    return WLSQ_OK;
}

Scratch &scratch_for(Device *d, hipStream_t st)
{
    auto &p = d->scratch[st];
    if (!p)
        p = std::make_unique<Scratch>();
    return *p;
}

// The kappa grid: nk samples of spacing dkappa. The lookup's rounding is
// exact up to 2^22 samples, and the fused kernel addresses table rows of L + 1
// entries with 32-bit byte offsets (L = 0: no table addressed here).
int check_kappa(int nk, float dkappa, int L)
{
    if (nk <= 0 || !(dkappa > 0.0f))
        return fail(WLSQ_EINVAL, "invalid kappa grid (nk=%d dkappa=%g)", nk, (double)dkappa);
    if (nk > kMaxKappaSamples)
        return fail(WLSQ_EUNSUPPORTED, "nk=%d above %d kappa samples", nk, kMaxKappaSamples);
    if (L > 0 && (unsigned long long)nk * (unsigned long long)(L + 1) * 8ull > 0xffffffffull)
        return fail(WLSQ_EUNSUPPORTED, "table of %d rows of %d taps above 4 GiB", nk, L);
    return WLSQ_OK;
}

int check_common(int ns, int nextrap, int nz, int nt, int nf, int nx, int M)
{
    if (ns < 0 || nextrap < 0 || nz < 0 || nt < 0 || nf < 0 || nx < 0 || M < 0)
        return fail(WLSQ_EINVAL, "negative size (ns=%d nextrap=%d nz=%d nt=%d nf=%d nx=%d M=%d)",
                    ns, nextrap, nz, nt, nf, nx, M);
    if (nz < nextrap)
        return fail(WLSQ_EINVAL, "nz (%d) must be >= nextrap (%d)", nz, nextrap);
    if (nt < nf)
        return fail(WLSQ_EINVAL, "nt (%d) must be >= nf (%d)", nt, nf);
    if (M > 1024)
        return fail(WLSQ_EUNSUPPORTED, "operator half-length M=%d above 1024", M);
    return WLSQ_OK;
}

// ------------------------------------------------------------------- plain path
// Runs nextrap steps for `nsh` shots with S shots per thread. src/dst fields
// have shot stride src_stride/dst_stride and row stride nx; img has shot
// stride img_stride and row stride nx.
int plain_sequence(Scratch &sc, int S, int nsh, int nextrap, int nf, int nx, int M,
                   const float2 *tapF, const float2 *tapB, const float2 *srcF,
                   const float2 *srcB, long long src_stride, float2 *dstF, float2 *dstB,
                   long long dst_stride, float *img, long long img_stride, hipStream_t st)
{
    if (nextrap <= 0 || nsh <= 0)
        return WLSQ_OK;
    const long long plane = (long long)nf * nx;
    const int G = plain_groups(nf);
    const long long part_stride = (long long)nextrap * G * nx;
    const int L = 2 * M + 1;
    const long long depth_slice = (long long)nf * L * nx;
    if (nf > 0) {
        for (int b = 0; b < 2; ++b) {
            WLSQ_TRY(sc.pingF[b].reserve(sizeof(float2) * (size_t)(nsh * plane)));
            WLSQ_TRY(sc.pingB[b].reserve(sizeof(float2) * (size_t)(nsh * plane)));
        }
        WLSQ_TRY(sc.part.reserve(sizeof(float) * (size_t)(nsh * part_stride)));
        for (int l = 0; l < nextrap; ++l) {
            const float2 *inF = l == 0 ? srcF : sc.pingF[(l - 1) & 1].as<float2>();
            const float2 *inB = l == 0 ? srcB : sc.pingB[(l - 1) & 1].as<float2>();
            const long long in_stride = l == 0 ? src_stride : plane;
            HIP_TRY(launch_plain_step(S, tapF + l * depth_slice, tapB + l * depth_slice, inF, inB,
                                      in_stride, sc.pingF[l & 1].as<float2>(),
                                      sc.pingB[l & 1].as<float2>(),
                                      sc.part.as<float>() + (long long)l * G * nx, part_stride, nsh,
                                      nf, nx, M, st));
        }
        const int last = (nextrap - 1) & 1;
        HIP_TRY(launch_copy_rows(sc.pingF[last].as<float2>(), plane, dstF, dst_stride, nsh, nf, nx, st));
        HIP_TRY(launch_copy_rows(sc.pingB[last].as<float2>(), plane, dstB, dst_stride, nsh, nf, nx, st));
        HIP_TRY(launch_plain_reduce(sc.part.as<float>(), img, nsh, img_stride, nextrap, G, nx, st));
    } else {
        // no frequencies: the image rows are sums over nothing
        for (int s = 0; s < nsh; ++s)
            HIP_TRY(launch_fill_f32(img + s * img_stride, (long long)nextrap * nx, 0.f, st));
    }
    return WLSQ_OK;
}

// shots_per_launch == 1: one launch sequence per shot (the reference's
// structure); otherwise all shots in each depth launch, S per thread.
int plain_run(Scratch &sc, int shots_per_launch, int ns, int nextrap, int nf, int nx, int M,
              const float2 *tapF, const float2 *tapB, const float2 *srcF, const float2 *srcB,
              long long src_stride, float2 *dstF, float2 *dstB, long long dst_stride, float *img,
              long long img_stride, hipStream_t st)
{
    if (shots_per_launch == 1) {
        for (int s = 0; s < ns; ++s)
            WLSQ_TRY(plain_sequence(sc, 1, 1, nextrap, nf, nx, M, tapF, tapB, srcF + s * src_stride,
                                    srcB + s * src_stride, src_stride, dstF + s * dst_stride,
                                    dstB + s * dst_stride, dst_stride, img + s * img_stride,
                                    img_stride, st));
        return WLSQ_OK;
    }
    return plain_sequence(sc, shots_per_launch, ns, nextrap, nf, nx, M, tapF, tapB, srcF, srcB,
                          src_stride, dstF, dstB, dst_stride, img, img_stride, st);
}

int transpose_tables(Scratch &sc, int nextrap, int nf, int nx, int M, const float2 *natF,
                     const float2 *natB, hipStream_t st)
{
    const int L = 2 * M + 1;
    const long long slices = (long long)nextrap * nf;
    const size_t bytes = sizeof(float2) * (size_t)slices * L * nx;
    WLSQ_TRY(sc.tapF.reserve(bytes));
    WLSQ_TRY(sc.tapB.reserve(bytes));
    // launch in runs of at most 65535 slices (grid y)
    for (long long s0 = 0; s0 < slices; s0 += 65535) {
        const long long n = std::min<long long>(65535, slices - s0);
        const long long off = s0 * L * nx;
        HIP_TRY(launch_transpose_tapmajor(natF + off, sc.tapF.as<float2>() + off, n, nx, L, st));
        HIP_TRY(launch_transpose_tapmajor(natB + off, sc.tapB.as<float2>() + off, n, nx, L, st));
    }
    return WLSQ_OK;
}

// ------------------------------------------------------------------- planner
struct PlanImpl {
    int M = 0;
    int L = 1;
    FusedShape shape{4, 1, 1, 256, false, false, false, false};
    size_t lds = 0;
    int img_lds = 1;
    int img_xpose = 0;
    int nplanes = 1;             // image planes of the direct atomics
    long long img_extent = 0;    // floats of the image the shots touch
    DevBuf planes;
    int exact_image = 0;         // image sums accumulated exactly, rounded once after the launch
    DevBuf xacc;                 // their accumulators: kExactPlanes planes of img_extent words
    int strip_gathers = 1;
    int double_rows = 0;
    int depth_steps = 0;         // depth steps per launch (0: all)
    int max_nz = 0;              // most depth steps a shot runs in one launch
    long long f_extent = 0, b_extent = 0;  // elements of the F and B pools the shots use
    DevBuf fields;               // copies of the pools for split launches without write-back
    long long points = 0;
    std::vector<Group> groups;
    std::vector<Item> items;
    Group *d_groups = nullptr;
    Item *d_items = nullptr;
    float2 *d_tt = nullptr;  // row-padded table copies: forward at 0, backward at Lp * tt_nk
    int tt_nk = 0;
    int device = 0;
    std::atomic<int> graphs{0};  // live graphs that captured d_tt: its size is then fixed
    ~PlanImpl()
    {
        if (d_groups)
            (void)hipFree(d_groups);
        if (d_items)
            (void)hipFree(d_items);
        if (d_tt)
            (void)hipFree(d_tt);
        if (fields.p)
            (void)hipFree(fields.p);
        if (planes.p)
            (void)hipFree(planes.p);
        if (xacc.p)
            (void)hipFree(xacc.p);
    }
    // table copy storage for nk samples (allocates: call outside graph capture).
    // A graph of this plan holds d_tt: it is never reallocated while one lives.
    int reserve_tables(int nk)
    {
        if (nk <= tt_nk)
            return WLSQ_OK;
        if (graphs.load() > 0)
            return fail(WLSQ_EINVAL,
                        "nk=%d above the %d kappa samples a captured graph of this plan holds "
                        "(destroy the graph or use a new plan)", nk, tt_nk);
        if (d_tt)
            (void)hipFree(d_tt);
        d_tt = nullptr;
        tt_nk = 0;
        hipError_t e = hipMalloc(&d_tt, sizeof(float2) * 2 * (size_t)(L + 1) * nk);
        if (e != hipSuccess)
            return fail(WLSQ_ENOMEM, "hipMalloc(table copies): %s", hipGetErrorString(e));
        tt_nk = nk;
        return WLSQ_OK;
    }
};

bool same_shape(const wlsq_shot &a, const wlsq_shot &b)
{
    return a.nx == b.nx && a.nz == b.nz && a.nf == b.nf && a.omega_off == b.omega_off &&
           a.slow_off == b.slow_off && a.slow_ld == b.slow_ld && a.ld == b.ld &&
           a.img_ld == b.img_ld && a.first_step == b.first_step;
}

// Depth steps shot `s` runs in a launch of a plan: every step (round_steps
// 0), or one round of a continuous batch, [first_step, first_step + round_steps)
// cut at nz.
long long shot_steps(const wlsq_shot &s, int round_steps)
{
    if (round_steps <= 0)
        return s.nz;
    return std::max(0ll, std::min<long long>(s.nz, (long long)s.first_step + round_steps) - s.first_step);
}

// Frequency rows of the tasks of a shot of nf rows, at most W per task, in
// frequency order. A task of w rows fills w * strips slots of the block's C
// passes; with `balance` the rows are split to minimise the summed depth-step
// time of the tasks, modelled as the strip passes of the busiest SIMD (a
// 1024-thread workgroup is alone on its CU and wavefront v runs on SIMD
// v mod 4; the others idle at the step's barriers) plus a quarter pass per
// task for its flush and image atomics. Example: 1500-wide rows (375 strips of
// 4 x) at C = 2 and W = 5 fill waves 0-15 and 0-13 (busiest SIMD 8 passes,
// 5 rows), where tasks of 4 rows fill 0-15 and 0-7 (6 passes, 4 rows): per row
// 1.6 against 1.5 passes. On MI355X, batches of one aperture at fixed rows per
// task give a step time of about a * (busiest SIMD's passes) + b with b near a
// (1500-wide rows: 7 % faster in tasks of 4 rows than of 5); on a batch of 128
// shots of apertures 256-2048 the split is worth 0.3 %.
std::vector<int> task_rows(int nf, int W, int strips, int block, int C, bool balance)
{
    std::vector<int> out;
    if (!balance || W <= 1 || W > 64) {
        for (int j0 = 0; j0 < nf; j0 += W)
            out.push_back(std::min(W, nf - j0));
        return out;
    }
    // cost[w] in quarter passes
    const int waves = block / 64;
    std::vector<long long> cost(W + 1, 0);
    for (int w = 1; w <= W; ++w) {
        int simd[4] = {0, 0, 0, 0};
        for (int v = 0; v < waves; ++v)
            for (int c = 0; c < C; ++c)
                if (v * 64 + c * block < w * strips)
                    ++simd[v & 3];
        cost[w] = 4ll * std::max(std::max(simd[0], simd[1]), std::max(simd[2], simd[3])) + 1;
    }
    std::vector<long long> best(nf + 1, -1);
    std::vector<int> last(nf + 1, 0);
    best[0] = 0;
    for (int n = 1; n <= nf; ++n) {
        for (int w = std::min(W, n); w >= 1; --w) {  // ties: the larger task
            const long long c = best[n - w] + cost[w];
            if (best[n] < 0 || c < best[n]) {
                best[n] = c;
                last[n] = w;
            }
        }
    }
    for (int n = nf; n > 0; n -= last[n])
        out.push_back(last[n]);
    std::reverse(out.begin(), out.end());
    return out;
}

// The work list and launch shape of a batch; with `upload` the group and
// task descriptors are copied to plan-owned device memory (the allocation-free
// entry places them itself). round_steps > 0: one round of a continuous batch,
// each shot running round_steps steps from its first_step (shots with nothing
// left have no tasks); round_steps 0: every first_step must be 0.
int build_plan(Device *d, const wlsq_shot *shots, int nshots, int M, const wlsq_tuning *tun,
               std::shared_ptr<PlanImpl> &out, bool upload = true, int round_steps = 0)
{
    if (nshots < 0 || (nshots > 0 && !shots))
        return fail(WLSQ_EINVAL, "shots array is NULL");
    if (M < 0 || M > 1024)
        return fail(WLSQ_EINVAL, "invalid M=%d", M);
    wlsq_tuning t{};
    if (tun)
        t = *tun;
    auto P = std::make_shared<PlanImpl>();
    P->M = M;
    P->L = 2 * M + 1;
    P->device = d->ordinal;

    int max_nx = 0;
    for (int i = 0; i < nshots; ++i) {
        const wlsq_shot &s = shots[i];
        if (s.nx < 0 || s.nz < 0 || s.nf < 0 || s.omega_off < 0 || s.f_off < 0 || s.b_off < 0 ||
            s.img_off < 0 || s.slow_off < 0 || s.first_step < 0)
            return fail(WLSQ_EINVAL, "shot %d: negative size or offset", i);
        if (round_steps <= 0 && s.first_step != 0)
            return fail(WLSQ_EINVAL, "shot %d: first_step %d (only rounds of a continuous batch "
                                     "start below the surface)", i, s.first_step);
        if (s.nx > 0 && s.nf > 0 && shot_steps(s, round_steps) > 0) {
            if (s.ld < s.nx || s.img_ld < s.nx || s.slow_ld < s.nx)
                return fail(WLSQ_EINVAL, "shot %d: row stride below aperture (ld=%d img_ld=%d slow_ld=%d nx=%d)",
                            i, s.ld, s.img_ld, s.slow_ld, s.nx);
            max_nx = std::max(max_nx, s.nx);
        }
        P->points += (long long)s.nx * shot_steps(s, round_steps) * s.nf;
    }

    // longest run of consecutive same-shaped shots (they can share a group)
    int best_run = 1;
    for (int i = 1, run = 1; i < nshots; ++i) {
        run = same_shape(shots[i], shots[i - 1]) ? run + 1 : 1;
        best_run = std::max(best_run, run);
    }
    // The CPU path's rounding is compiled for one shot per group, 4 x per
    // thread and 1024 threads, taps summed in order.
    if (t.rounding < 0 || t.rounding > 1)
        return fail(WLSQ_EINVAL, "rounding must be 0 or 1 (got %d)", t.rounding);
    const bool exact = t.rounding == 1;
    if (exact) {
        if ((t.x_per_thread != 0 && t.x_per_thread != 4) ||
            (t.shots_per_group != 0 && t.shots_per_group != 1) ||
            (t.block != 0 && t.block <= 512))
            return fail(WLSQ_EINVAL, "rounding=1 needs 4 x per thread, one shot per group and "
                                     "more than 512 threads");
        t.x_per_thread = 4;
        t.shots_per_group = 1;
        t.tables_symmetric = 0;
    }
    // x per thread and shots per group; the defaults follow an autotuner
    // sweep on MI355X over apertures 512-8192 and 1-16 shots per group
    int R = t.x_per_thread;
    if (R == 0)
        R = (best_run >= 2 && max_nx <= 1024 && t.shots_per_group != 1) ? 2 : 4;
    if (R != 2 && R != 4 && R != 8)
        return fail(WLSQ_EINVAL, "x_per_thread must be 2, 4 or 8 (got %d)", R);
    int S = t.shots_per_group;
    if (S == 0)
        S = R == 8 ? 1 : R == 2 ? (best_run >= 8 ? 4 : 2) : (best_run >= 2 && max_nx <= 2048 ? 2 : 1);
    if (S != 1 && S != 2 && S != 4)
        return fail(WLSQ_EINVAL, "shots_per_group must be 1, 2 or 4 (got %d)", S);
    if (R == 8 && S != 1)
        return fail(WLSQ_EUNSUPPORTED, "8 x per thread supports 1 shot per group");
    if (R == 2 && S == 1)
        return fail(WLSQ_EUNSUPPORTED, "2 x per thread supports 2 or 4 shots per group");
    // Single-shot groups of 4 x per thread and two-shot groups of 2 x also run
    // 1024-thread workgroups (their registers fit four waves per SIMD); the
    // other shapes run up to 512.
    const int max_block = ((R == 4 && S == 1) || (R == 2 && S == 2)) ? 1024 : 512;
    const int max_strips = (max_nx + R - 1) / R;
    // strips per thread: as tuned, else 2 when the aperture has more strips
    // than the workgroup has threads (8 x per thread, or 4 with 1024 threads)
    const int threads = t.block > 0 ? t.block : max_block;
    if (t.strips_per_thread < 0 || t.strips_per_thread > 2)
        return fail(WLSQ_EINVAL, "strips_per_thread must be 0, 1 or 2 (got %d)", t.strips_per_thread);
    const int C = t.strips_per_thread > 0 ? t.strips_per_thread : (max_strips <= threads ? 1 : 2);
    if (C == 2 && !((R == 4 && S == 1) || R == 8))
        return fail(WLSQ_EUNSUPPORTED, "two strips per thread are compiled for 4 x per thread with one "
                                       "shot per group and for 8 x per thread");
    if (C == 1 && max_strips > threads && t.strips_per_thread == 1)
        return fail(WLSQ_EUNSUPPORTED, "aperture nx=%d needs %d strips: above %d threads with one strip "
                                       "per thread", max_nx, max_strips, threads);
    if (max_strips > (R == 2 ? 1 : C) * max_block || (C == 2 && R == 2) || (C == 2 && R == 4 && S != 1))
        return fail(WLSQ_EUNSUPPORTED, "aperture nx=%d needs %d strips of %d x (max %d)", max_nx,
                    max_strips, R, (R == 2 || S != 1 ? 1 : 2) * max_block);
    if (C == 2 && R == 4 && t.block > 0 && t.block <= 512)
        return fail(WLSQ_EINVAL, "aperture nx=%d with 4 x per thread needs block > 512 (got %d)",
                    max_nx, t.block);

    int block = t.block;
    if (block == 0) {
        block = (((max_strips + C - 1) / C + 63) / 64) * 64;
        // single-shot groups: 1024 threads (several frequencies per task on
        // narrow apertures); 4 x per thread with shots: 512; else >= 256
        block = std::max(block, (R == 4 && S == 1) ? 1024 : R == 4 ? 512 : 256);
    }
    if (block < 64 || block > max_block || block % 64 || block * C < max_strips ||
        (C == 2 && R == 4 && block <= 512))
        return fail(WLSQ_EINVAL, "block=%d invalid for %d strips (multiple of 64, <= %d)", block,
                    max_strips, max_block);

    const size_t lds_max = std::max((size_t)d->props.sharedMemPerBlock,
                                    (size_t)d->props.sharedMemPerBlockOptin);
    size_t budget = t.lds_budget > 0 ? (size_t)t.lds_budget : std::min<size_t>(lds_max, 80 * 1024);
    if (budget > lds_max)
        budget = lds_max;

    // groups: runs of up to S same-shaped consecutive shots
    std::vector<std::pair<int, int>> runs;  // (first shot, count)
    for (int i = 0; i < nshots;) {
        const wlsq_shot &s = shots[i];
        if (s.nx == 0 || s.nf == 0 || shot_steps(s, round_steps) == 0) {
            ++i;
            continue;
        }
        int n = 1;
        while (n < S && i + n < nshots && same_shape(shots[i + n], s))
            ++n;
        runs.emplace_back(i, n);
        i += n;
    }
    struct Pending {
        Item item;
        long long cost;
    };
    auto run_shared = [&](const std::pair<int, int> &rn) {
        for (int k = 1; k < rn.second; ++k)
            if (shots[rn.first + k].img_off != shots[rn.first].img_off)
                return false;
        return true;
    };
    // The image sum over a task's frequencies is formed in LDS: straight from
    // the new field rows after each step's barrier (mode 4, the default for
    // one shot per group: no partial rows, so more frequency rows fit a task,
    // and lane-contiguous atomics for every aperture - on 8192-wide survey
    // shots 25 % faster than mode 3), or from per-frequency partial rows
    // (mode 1, the default for groups of shots) unless some group cannot fit
    // them at W = 1; then direct atomics, made lane-contiguous through a
    // wave-private LDS segment in single-frequency tasks (mode 3), or per
    // point as computed (mode 2).
    if (t.image_mode < 0 || t.image_mode > 4)
        return fail(WLSQ_EINVAL, "image_mode must be 0..4 (got %d)", t.image_mode);
    const bool rows_img = t.image_mode == 4 || (t.image_mode == 0 && S == 1);
    int img_lds = (t.image_mode == 2 || t.image_mode == 3) ? 0 : 1;
    for (auto &rn : runs) {
        if (!img_lds || rows_img)
            break;
        const int rows = run_shared(rn) ? 1 : S;
        if (fused_lds_bytes(shots[rn.first].nx, M, R, S, 1, rows) > lds_max) {
            if (t.image_mode == 1)
                return fail(WLSQ_EUNSUPPORTED, "image rows do not fit in LDS for nx=%d",
                            shots[rn.first].nx);
            img_lds = 0;
        }
    }
    if (t.image_mode == 4 && S != 1)
        return fail(WLSQ_EUNSUPPORTED, "image_mode 4 is compiled for one shot per group");
    if (rows_img)
        img_lds = 2;
    std::vector<Pending> pend;
    // the task rows of each (frequencies, rows per task, strips) the plan meets
    // (a batch of many shots of one shape solves the split once)
    std::map<std::tuple<int, int, int>, std::vector<int>> rows_memo;
    size_t lds_need = 0;
    for (auto &rn : runs) {
        const wlsq_shot &s = shots[rn.first];
        Group g{};
        for (int k = 0; k < kMaxShotsPerGroup; ++k) {
            const wlsq_shot &sk = shots[rn.first + std::min(k, rn.second - 1)];
            g.f_off[k] = sk.f_off;
            g.b_off[k] = sk.b_off;
            g.img_off[k] = sk.img_off;
        }
        g.slow_off = s.slow_off;
        g.nshots = rn.second;
        g.ld = s.ld;
        g.img_ld = s.img_ld;
        g.slow_ld = s.slow_ld;
        g.nx = s.nx;
        g.nz = s.nz;
        g.nf = s.nf;
        g.omega_off = s.omega_off;
        g.img_shared = run_shared(rn) ? 1 : 0;
        g.z0 = s.first_step;
        const int img_rows = img_lds == 1 ? (g.img_shared ? 1 : S) : 0;
        const int strips = (s.nx + R - 1) / R;
        int W = block * C / strips;
        if (t.omegas_per_item > 0)
            W = std::min(W, t.omegas_per_item);
        W = std::min(W, s.nf);
        while (W > 1 && fused_lds_bytes(s.nx, M, R, S, W, img_rows) > budget)
            --W;
        if (W < 1)
            W = 1;
        const size_t need = fused_lds_bytes(s.nx, M, R, S, W, img_rows);
        if (need > lds_max)
            return fail(WLSQ_EUNSUPPORTED,
                        "aperture nx=%d with M=%d needs %zu bytes of LDS (device max %zu)", s.nx,
                        M, need, lds_max);
        lds_need = std::max(lds_need, need);
        const int gi = (int)P->groups.size();
        P->groups.push_back(g);
        auto memo = rows_memo.find(std::make_tuple(s.nf, W, strips));
        if (memo == rows_memo.end())
            memo = rows_memo
                       .emplace(std::make_tuple(s.nf, W, strips),
                                task_rows(s.nf, W, strips, block, C,
                                          t.omegas_per_item == 0 && block == 1024))
                       .first;
        int j0 = 0;
        for (const int nw : memo->second) {
            Pending pd;
            pd.item.group = gi;
            pd.item.j0 = j0;
            pd.item.nw = nw;
            pd.item.cost_rank = 0;
            pd.cost = shot_steps(s, round_steps) * s.nx * pd.item.nw * rn.second;
            pend.push_back(pd);
            j0 += nw;
        }
    }
    // longest first: the hardware dispatches workgroups in index order
    std::stable_sort(pend.begin(), pend.end(),
                     [](const Pending &a, const Pending &b) { return a.cost > b.cost; });
    for (size_t i = 0; i < pend.size(); ++i) {
        pend[i].item.cost_rank = (int)i;
        P->items.push_back(pend[i].item);
    }
    if (t.tap_loads < 0 || t.tap_loads > 2)
        return fail(WLSQ_EINVAL, "tap_loads must be 0, 1 or 2 (got %d)", t.tap_loads);
    // paired tap loads relieve the address unit when no shots share the gathers
    // (1024-thread workgroups too, now that their tap loop is fenced: 4-14 %)
    const bool paired = t.tap_loads == 2 || (t.tap_loads == 0 && S == 1 && R <= 4);
    P->shape = FusedShape{R, S, C, block, t.tables_conjugate != 0, t.tables_symmetric != 0, paired,
                          exact};
    P->img_lds = img_lds;
    // the transpose and the image planes are compiled for the one-shot
    // 1024-thread shapes (the default for apertures whose partial rows do not
    // fit); other shapes add each point into the image
    const bool xp_shape = R == 4 && S == 1 && block > 512;
    P->img_xpose = !img_lds && t.image_mode != 2 && xp_shape;
    // the transpose segments: 64 x R floats per wavefront, one copy
    const size_t seg_bytes = P->img_xpose ? sizeof(float) * (size_t)block * R : 0;
    if (t.strip_gathers < 0 || t.strip_gathers > 2)
        return fail(WLSQ_EINVAL, "strip_gathers must be 0, 1 or 2 (got %d)", t.strip_gathers);
    P->strip_gathers = t.strip_gathers == 2 ? 0 : 1;
    // Two copies of the rows (one barrier per step) when every task's doubled
    // footprint fits the device's LDS.
    if (t.double_rows < 0 || t.double_rows > 2)
        return fail(WLSQ_EINVAL, "double_rows must be 0, 1 or 2 (got %d)", t.double_rows);
    if (t.double_rows != 2 && C == 1 && 2 * lds_need + seg_bytes <= lds_max) {
        P->double_rows = 1;
        lds_need *= 2;
    } else if (t.double_rows == 1) {
        return fail(WLSQ_EUNSUPPORTED, "two copies of the rows need %zu bytes of LDS (device max %zu)",
                    2 * lds_need + seg_bytes, lds_max);
    }
    lds_need += seg_bytes;
    if (lds_need > lds_max)
        return fail(WLSQ_EUNSUPPORTED, "rows and transpose segments need %zu bytes of LDS (device max %zu)",
                    lds_need, lds_max);
    P->lds = lds_need;
    if (t.depth_steps < 0)
        return fail(WLSQ_EINVAL, "depth_steps must be >= 0 (got %d)", t.depth_steps);
    P->depth_steps = t.depth_steps;
    for (int i = 0; i < nshots; ++i) {
        const wlsq_shot &sh = shots[i];
        if (sh.nx == 0 || sh.nf == 0 || shot_steps(sh, round_steps) == 0)
            continue;
        P->max_nz = std::max(P->max_nz, (int)shot_steps(sh, round_steps));
        P->f_extent = std::max(P->f_extent, sh.f_off + (long long)(sh.nf - 1) * sh.ld + sh.nx);
        P->b_extent = std::max(P->b_extent, sh.b_off + (long long)(sh.nf - 1) * sh.ld + sh.nx);
        P->img_extent = std::max(P->img_extent, sh.img_off + (long long)(sh.nz - 1) * sh.img_ld + sh.nx);
    }
    // Image planes for the direct atomics: auto = 8 while they take at most
    // 512 MiB, halved until they fit (none below 2).
    if (t.image_planes < 0 || t.image_planes > kMaxImagePlanes)
        return fail(WLSQ_EINVAL, "image_planes must be 0..%d (got %d)", kMaxImagePlanes,
                    t.image_planes);
    if (P->img_lds == 2 && t.image_planes >= 2) {
        P->nplanes = t.image_planes;  // sums from the rows into planes: on request only
    } else if (P->img_xpose) {
        int K = t.image_planes;
        if (K == 0) {
            K = 8;
            while (K >= 2 && (double)K * (double)P->img_extent * sizeof(float) > (double)(1ll << 29))
                K /= 2;
        }
        P->nplanes = K >= 2 ? K : 1;
    }
    // Exact image sums: integer accumulation of each task's flushed sums (the
    // flushes of image modes 1 and 4), rounded once after the launch.
    if (t.exact_image < 0 || t.exact_image > 1)
        return fail(WLSQ_EINVAL, "exact_image must be 0 or 1 (got %d)", t.exact_image);
    if (t.exact_image) {
        if (P->img_lds == 0)
            return fail(WLSQ_EUNSUPPORTED, "exact_image needs image mode 1 or 4 (the task sums a "
                                           "flush adds); these shapes add per point");
        if (!P->shape.conj)
            return fail(WLSQ_EUNSUPPORTED, "exact_image is compiled for conjugate tables "
                                           "(tables_conjugate = 1)");
        P->exact_image = 1;
        P->nplanes = 1;
    }
    if (upload && !P->groups.empty()) {
        HIP_TRY(hipMalloc(&P->d_groups, sizeof(Group) * P->groups.size()));
        HIP_TRY(hipMemcpy(P->d_groups, P->groups.data(), sizeof(Group) * P->groups.size(),
                          hipMemcpyHostToDevice));
    }
    if (upload && !P->items.empty()) {
        HIP_TRY(hipMalloc(&P->d_items, sizeof(Item) * P->items.size()));
        HIP_TRY(hipMemcpy(P->d_items, P->items.data(), sizeof(Item) * P->items.size(),
                          hipMemcpyHostToDevice));
    }
    out = P;
    return WLSQ_OK;
}

// Kernel arguments of a plan's launch over every depth step, image planes off.
FusedArgs fused_args(const PlanImpl &P, const Item *items, const Group *groups, const float2 *ttF,
                     const float2 *ttB, const float *slowness, const float *omega, float kappa0,
                     float dkappa, int nk, wlsq_c32 *forw, wlsq_c32 *back, float *image,
                     int write_back)
{
    FusedArgs a{};
    a.items = items;
    a.groups = groups;
    a.F = reinterpret_cast<float2 *>(forw);
    a.B = reinterpret_cast<float2 *>(back);
    a.img = image;
    a.omega = omega;
    a.slow = slowness;
    a.TF = ttF;
    a.TB = P.shape.conj ? ttF : ttB;
    a.nk = nk;
    a.kappa0 = kappa0;
    a.inv_dk = inv_dkappa(dkappa);
    a.L = P.L;
    a.write_back = write_back;
    a.img_lds = P.img_lds;
    a.img_xpose = P.img_xpose;
    a.nplanes = 1;
    a.strip_gathers = P.strip_gathers;
    a.double_rows = P.double_rows;
    a.l0 = 0;
    a.l1 = P.max_nz;
    return a;
}

int launch_plan(PlanImpl &P, const float *slowness, const float *omega, float kappa0,
                float dkappa, int nk, const wlsq_c32 *tf, const wlsq_c32 *tb, wlsq_c32 *forw,
                wlsq_c32 *back, float *image, int write_back, hipStream_t st)
{
    if (P.items.empty())
        return WLSQ_OK;
    if (!slowness || !omega || !tf || (!tb && !P.shape.conj) || !forw || !back || !image)
        return fail(WLSQ_EINVAL, "NULL device pointer");
    WLSQ_TRY(check_kappa(nk, dkappa, P.L));
    WLSQ_TRY(P.reserve_tables(nk));
    float2 *ttF = P.d_tt;
    float2 *ttB = P.d_tt + (size_t)(P.L + 1) * nk;
    HIP_TRY(launch_table_pad(reinterpret_cast<const float2 *>(tf), ttF, nk, P.L, st));
    if (!P.shape.conj)
        HIP_TRY(launch_table_pad(reinterpret_cast<const float2 *>(tb), ttB, nk, P.L, st));
    FusedArgs a = fused_args(P, P.d_items, P.d_groups, ttF, ttB, slowness, omega, kappa0, dkappa,
                             nk, forw, back, image, write_back);
    if (P.nplanes > 1) {
        WLSQ_TRY(P.planes.reserve(sizeof(float) * (size_t)P.nplanes * (size_t)P.img_extent));
        a.planes = P.planes.as<float>();
        a.plane_stride = P.img_extent;
        a.nplanes = P.nplanes;
        HIP_TRY(launch_fill_f32(a.planes, (long long)P.nplanes * P.img_extent, 0.f, st));
    }
    if (P.exact_image) {
        // zeroed by a fill kernel (no memset node when captured in a graph)
        WLSQ_TRY(P.xacc.reserve(sizeof(unsigned long long) * kExactPlanes * (size_t)P.img_extent));
        a.xacc = P.xacc.as<unsigned long long>();
        a.xstride = P.img_extent;
        HIP_TRY(launch_fill_f32(P.xacc.as<float>(), 2ll * kExactPlanes * P.img_extent, 0.f, st));
    }
    const int ds = P.depth_steps;
    if (ds == 0 || ds >= P.max_nz) {
        a.l0 = 0;
        a.l1 = P.max_nz;
        HIP_TRY(launch_fused(P.shape, (int)P.items.size(), P.lds, st, a));
        if (P.nplanes > 1)
            HIP_TRY(launch_planes_reduce(a.planes, a.plane_stride, a.nplanes, image, P.img_extent, st));
        if (P.exact_image)
            HIP_TRY(launch_exact_finish(a.xacc, a.xstride, P.img_extent, image, st));
        return WLSQ_OK;
    }
    // Depth range split over launches of ds steps: each launch starts from
    // the fields the previous one wrote back. Without write-back the caller's
    // fields stay untouched: the launches run on copies.
    if (!write_back) {
        WLSQ_TRY(P.fields.reserve(sizeof(float2) * (size_t)(P.f_extent + P.b_extent)));
        float2 *cf = P.fields.as<float2>();
        float2 *cb = cf + P.f_extent;
        HIP_TRY(hipMemcpyAsync(cf, forw, sizeof(float2) * (size_t)P.f_extent, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(cb, back, sizeof(float2) * (size_t)P.b_extent, hipMemcpyDeviceToDevice, st));
        a.F = cf;
        a.B = cb;
    }
    a.write_back = 1;
    for (int l0 = 0; l0 < P.max_nz; l0 += ds) {
        a.l0 = l0;
        a.l1 = l0 + ds;
        HIP_TRY(launch_fused(P.shape, (int)P.items.size(), P.lds, st, a));
    }
    if (P.nplanes > 1)
        HIP_TRY(launch_planes_reduce(a.planes, a.plane_stride, a.nplanes, image, P.img_extent, st));
    if (P.exact_image)
        HIP_TRY(launch_exact_finish(a.xacc, a.xstride, P.img_extent, image, st));
    return WLSQ_OK;
}

std::vector<wlsq_shot> uniform_shots(int ns, int nextrap, int nz, int nt, int nf, int nx,
                                     int slow_ld)
{
    std::vector<wlsq_shot> v((size_t)ns);
    for (int s = 0; s < ns; ++s) {
        wlsq_shot &d = v[(size_t)s];
        d = wlsq_shot{};
        d.f_off = (long long)s * nt * nx;
        d.b_off = d.f_off;
        d.img_off = (long long)s * nz * nx;
        d.slow_off = 0;
        d.ld = nx;
        d.img_ld = nx;
        d.slow_ld = slow_ld;
        d.nx = nx;
        d.nz = nextrap;
        d.nf = nf;
        d.omega_off = 0;
    }
    return v;
}

int ktable_device_impl(Device *d, int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                       const float *slowness, int slow_ld, const float *omega, float kappa0,
                       float dkappa, int nk, const wlsq_c32 *tf, wlsq_c32 *forw,
                       const wlsq_c32 *tb, wlsq_c32 *back, float *image, int write_back,
                       const wlsq_tuning *tuning, hipStream_t st)
{
    WLSQ_TRY(check_common(ns, nextrap, nz, nt, nf, nx, M));
    if (ns == 0 || nextrap == 0 || nx == 0)
        return WLSQ_OK;
    if (slow_ld < nx)
        return fail(WLSQ_EINVAL, "slow_ld (%d) < nx (%d)", slow_ld, nx);
    // image rows l < nextrap are outputs: clear them, then accumulate
    for (int s = 0; s < ns; ++s)
        HIP_TRY(launch_fill_f32(image + (long long)s * nz * nx, (long long)nextrap * nx, 0.f, st));
    if (nf == 0)
        return WLSQ_OK;
    wlsq_tuning t{};
    if (tuning)
        t = *tuning;
    std::vector<long long> key{ns, nextrap, nz, nt, nf, nx, M, slow_ld, t.shots_per_group,
                               t.omegas_per_item, t.block, t.x_per_thread, t.lds_budget,
                               t.tables_conjugate, t.image_mode, t.tables_symmetric, t.tap_loads,
                               t.strip_gathers, t.double_rows, t.depth_steps, t.image_planes,
                               t.rounding, (long long)(intptr_t)st};  // a plan's table copies are per stream
    std::shared_ptr<PlanImpl> P;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        auto it = d->ktable_plans.find(key);
        if (it != d->ktable_plans.end())
            P = it->second;
    }
    if (!P) {
        auto shots = uniform_shots(ns, nextrap, nz, nt, nf, nx, slow_ld);
        WLSQ_TRY(build_plan(d, shots.data(), ns, M, &t, P));
        std::lock_guard<std::mutex> lk(d->mu);
        if (d->ktable_plans.size() > 64)
            d->ktable_plans.clear();
        d->ktable_plans[key] = P;
    }
    return launch_plan(*P, slowness, omega, kappa0, dkappa, nk, tf, tb, forw, back, image,
                       write_back, st);
}

bool host_tables_symmetric(const wlsq_c32 *t, int nk, int L)
{
    for (long long q = 0; q < nk; ++q)
        for (int k = 0; k < L / 2; ++k)
            if (std::memcmp(&t[q * L + k], &t[q * L + L - 1 - k], sizeof(wlsq_c32)) != 0)
                return false;
    return true;
}

bool host_tables_conjugate(const wlsq_c32 *tf, const wlsq_c32 *tb, long long n)
{
    for (long long i = 0; i < n; ++i) {
        uint32_t a, b, c, e;
        std::memcpy(&a, &tf[i].re, 4);
        std::memcpy(&b, &tb[i].re, 4);
        const float ni = -tf[i].im;
        std::memcpy(&c, &ni, 4);
        std::memcpy(&e, &tb[i].im, 4);
        if (a != b || c != e)
            return false;
    }
    return true;
}

}  // namespace

// ------------------------------------------------- launches into a given workspace
namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Layout of a plan's device data in one workspace: groups | tasks | forward
// table copy | backward table copy (when the tables are not conjugate).
struct WsLayout {
    size_t g_bytes = 0, i_bytes = 0, t_bytes = 0;
    size_t off_items = 0, off_tf = 0, off_tb = 0, need = 0;
};

WsLayout ws_layout(const PlanImpl &P, int nk)
{
    WsLayout w;
    w.g_bytes = sizeof(Group) * P.groups.size();
    w.i_bytes = sizeof(Item) * P.items.size();
    w.t_bytes = sizeof(float2) * (size_t)(P.L + 1) * (size_t)(nk > 0 ? nk : 0);
    w.off_items = align256(w.g_bytes);
    w.off_tf = w.off_items + align256(w.i_bytes);
    w.off_tb = w.off_tf + align256(w.t_bytes);
    w.need = w.off_tb + (P.shape.conj ? 0 : align256(w.t_bytes));
    return w;
}

// The plan's launch over every depth step with its descriptors and table
// copies placed in `ws` (layout ws_layout): kernels only, nothing allocated.
int launch_in_workspace(const PlanImpl &P, unsigned char *ws, const WsLayout &w,
                        const float *slowness, const float *omega, float kappa0, float dkappa,
                        int nk, const wlsq_c32 *tf, const wlsq_c32 *tb, wlsq_c32 *forw,
                        wlsq_c32 *back, float *image, int write_back, hipStream_t st)
{
    if (P.items.empty())
        return WLSQ_OK;
    Group *groups = reinterpret_cast<Group *>(ws);
    Item *items = reinterpret_cast<Item *>(ws + w.off_items);
    float2 *ttF = reinterpret_cast<float2 *>(ws + w.off_tf);
    float2 *ttB = reinterpret_cast<float2 *>(ws + w.off_tb);
    HIP_TRY(launch_upload(groups, P.groups.data(), w.g_bytes, st));
    HIP_TRY(launch_upload(items, P.items.data(), w.i_bytes, st));
    HIP_TRY(launch_table_pad(reinterpret_cast<const float2 *>(tf), ttF, nk, P.L, st));
    if (!P.shape.conj)
        HIP_TRY(launch_table_pad(reinterpret_cast<const float2 *>(tb), ttB, nk, P.L, st));
    const FusedArgs a = fused_args(P, items, groups, ttF, ttB, slowness, omega, kappa0, dkappa,
                                   nk, forw, back, image, write_back);
    HIP_TRY(launch_fused(P.shape, (int)P.items.size(), P.lds, st, a));
    return WLSQ_OK;
}

// Tuning of the workspace entries: every depth step in one launch, direct
// image atomics without planes.
int workspace_tuning(const wlsq_tuning *tuning, wlsq_tuning *t)
{
    *t = wlsq_tuning{};
    if (tuning)
        *t = *tuning;
    if (t->depth_steps != 0)
        return fail(WLSQ_EINVAL, "launches into a workspace run every depth step in one launch "
                                 "(depth_steps must be 0, got %d)", t->depth_steps);
    if (t->exact_image != 0)
        return fail(WLSQ_EINVAL, "exact_image needs a plan's accumulators (use a plan, "
                                 "wlsq_batch_device or wlsq_ktable_device)");
    t->image_planes = 1;
    return WLSQ_OK;
}

}  // namespace

// ------------------------------------------------------- allocation-free entry
// Device memory of wlsq_ktable_device_noalloc: the plan's group and task
// descriptors and the row-padded table copies. Each GPU the library runs on
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives for the current device.
constexpr size_t kNoAllocScratchBytes = 16ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// This device's copy of `scratch`, looked up on the first call per device.
int device_scratch(Device *d, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(d->mu);
    if (!d->scratch_ptr) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(WLSQ_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        d->scratch_ptr = ptr;
    }
    *out = static_cast<unsigned char *>(d->scratch_ptr);
    return WLSQ_OK;
}

}  // namespace

struct wlsq_plan_s {
    std::shared_ptr<PlanImpl> impl;
};

struct wlsq_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::shared_ptr<PlanImpl> plan;
};

int wlsq::set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" {

const char *wlsq_last_error(void) { return g_err.c_str(); }

int wlsq_version(void) { return 1; }

int wlsq_set_device(int device)
{
    HIP_TRY(hipSetDevice(device));
    return WLSQ_OK;
}

int extrapolate(int ns, int nextrap, int nz, int nt, int nf, int nx, int M, wlsq_c32 *w_op_forw,
                wlsq_c32 *forw_pulse, wlsq_c32 *w_op_back, wlsq_c32 *back_pulse, float *image)
{
    WLSQ_TRY(check_common(ns, nextrap, nz, nt, nf, nx, M));
    if (ns == 0 || nextrap == 0 || nx == 0)
        return WLSQ_OK;
    if (!w_op_forw || !w_op_back || !forw_pulse || !back_pulse || !image)
        return fail(WLSQ_EINVAL, "NULL array");
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    std::lock_guard<std::mutex> call_lk(d->host_call_mu);
    hipStream_t st = nullptr;
    WLSQ_TRY(host_stream(d, &st));
    Scratch *scp;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        scp = &scratch_for(d, st);
    }
    Scratch &sc = *scp;
    const int L = 2 * M + 1;
    const long long slice = (long long)nf * nx * L;  // one depth of one table
    const long long plane = (long long)nf * nx;

    if (nf > 0) {
        // tables: natural layout uploaded in chunks of depths, transposed to
        // tap-major on the device (replaces cpp_src/revOp.c:4-30 on the host)
        const size_t tab_bytes = sizeof(float2) * (size_t)(nextrap * slice);
        WLSQ_TRY(sc.tapF.reserve(tab_bytes));
        WLSQ_TRY(sc.tapB.reserve(tab_bytes));
        const long long chunk_depths =
            std::max<long long>(1, (64ll << 20) / (long long)(sizeof(float2) * slice));
        const long long cd = std::min<long long>(chunk_depths, nextrap);
        WLSQ_TRY(sc.stage.reserve(sizeof(float2) * (size_t)(cd * slice)));
        for (int pass = 0; pass < 2; ++pass) {
            const wlsq_c32 *src = pass == 0 ? w_op_forw : w_op_back;
            float2 *dst = pass == 0 ? sc.tapF.as<float2>() : sc.tapB.as<float2>();
            for (long long l0 = 0; l0 < nextrap; l0 += cd) {
                const long long n = std::min<long long>(cd, nextrap - l0);
                HIP_TRY(hipMemcpyAsync(sc.stage.p, src + l0 * slice, sizeof(float2) * (size_t)(n * slice),
                                       hipMemcpyHostToDevice, st));
                for (long long s0 = 0; s0 < n * nf; s0 += 65534) {
                    const long long m = std::min<long long>(65534, n * nf - s0);
                    HIP_TRY(launch_transpose_tapmajor(sc.stage.as<float2>() + s0 * nx * L,
                                                      dst + l0 * slice + s0 * nx * L, m, nx, L, st));
                }
            }
        }
        // fields: rows j < nf of each shot
        WLSQ_TRY(sc.fields_f.reserve(sizeof(float2) * (size_t)(ns * plane)));
        WLSQ_TRY(sc.fields_b.reserve(sizeof(float2) * (size_t)(ns * plane)));
        const size_t row_bytes = sizeof(float2) * (size_t)plane;
        const size_t shot_pitch = sizeof(float2) * (size_t)nt * nx;
        HIP_TRY(hipMemcpy2DAsync(sc.fields_f.p, row_bytes, forw_pulse, shot_pitch, row_bytes, ns,
                                 hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpy2DAsync(sc.fields_b.p, row_bytes, back_pulse, shot_pitch, row_bytes, ns,
                                 hipMemcpyHostToDevice, st));
    }
    WLSQ_TRY(sc.img.reserve(sizeof(float) * (size_t)ns * nextrap * nx));
    const int spl = ns >= 4 ? 4 : ns >= 2 ? 2 : 1;
    WLSQ_TRY(plain_run(sc, spl, ns, nextrap, nf, nx, M, sc.tapF.as<float2>(), sc.tapB.as<float2>(),
                       sc.fields_f.as<float2>(), sc.fields_b.as<float2>(), plane,
                       sc.fields_f.as<float2>(), sc.fields_b.as<float2>(), plane,
                       sc.img.as<float>(), (long long)nextrap * nx, st));
    if (nf > 0) {
        const size_t row_bytes = sizeof(float2) * (size_t)plane;
        const size_t shot_pitch = sizeof(float2) * (size_t)nt * nx;
        HIP_TRY(hipMemcpy2DAsync(forw_pulse, shot_pitch, sc.fields_f.p, row_bytes, row_bytes, ns,
                                 hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpy2DAsync(back_pulse, shot_pitch, sc.fields_b.p, row_bytes, row_bytes, ns,
                                 hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpy2DAsync(image, sizeof(float) * (size_t)nz * nx, sc.img.p,
                             sizeof(float) * (size_t)nextrap * nx, sizeof(float) * (size_t)nextrap * nx,
                             ns, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WLSQ_OK;
}

int wlsq_plain_device(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                      const wlsq_c32 *op_forw, wlsq_c32 *forw_pulse, const wlsq_c32 *op_back,
                      wlsq_c32 *back_pulse, float *image, int layout, int shots_per_launch,
                      void *stream)
{
    WLSQ_TRY(check_common(ns, nextrap, nz, nt, nf, nx, M));
    if (layout != WLSQ_LAYOUT_NATURAL && layout != WLSQ_LAYOUT_TAPMAJOR)
        return fail(WLSQ_EINVAL, "unknown table layout %d", layout);
    if (shots_per_launch != 1 && shots_per_launch != 2 && shots_per_launch != 4)
        return fail(WLSQ_EINVAL, "shots_per_launch must be 1, 2 or 4 (got %d)", shots_per_launch);
    if (ns == 0 || nextrap == 0 || nx == 0)
        return WLSQ_OK;
    if (!op_forw || !op_back || !forw_pulse || !back_pulse || !image)
        return fail(WLSQ_EINVAL, "NULL device pointer");
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(d->mu);
    Scratch &sc = scratch_for(d, st);
    const float2 *tapF = reinterpret_cast<const float2 *>(op_forw);
    const float2 *tapB = reinterpret_cast<const float2 *>(op_back);
    if (layout == WLSQ_LAYOUT_NATURAL && nf > 0) {
        WLSQ_TRY(transpose_tables(sc, nextrap, nf, nx, M, tapF, tapB, st));
        tapF = sc.tapF.as<float2>();
        tapB = sc.tapB.as<float2>();
    }
    const long long stride = (long long)nt * nx;
    return plain_run(sc, shots_per_launch, ns, nextrap, nf, nx, M, tapF, tapB,
                     reinterpret_cast<const float2 *>(forw_pulse),
                     reinterpret_cast<const float2 *>(back_pulse), stride,
                     reinterpret_cast<float2 *>(forw_pulse), reinterpret_cast<float2 *>(back_pulse),
                     stride, image, (long long)nz * nx, st);
}

int wlsq_ktable_device(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                       const float *slowness, int slow_ld, const float *omega, float kappa0,
                       float dkappa, int nk, const wlsq_c32 *table_forw, wlsq_c32 *forw_pulse,
                       const wlsq_c32 *table_back, wlsq_c32 *back_pulse, float *image,
                       int write_back, const wlsq_tuning *tuning, void *stream)
{
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    return ktable_device_impl(d, ns, nextrap, nz, nt, nf, nx, M, slowness, slow_ld, omega, kappa0,
                              dkappa, nk, table_forw, forw_pulse, table_back, back_pulse, image,
                              write_back, tuning, static_cast<hipStream_t>(stream));
}

int wlsq_ktable_device_noalloc(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                               const float *slowness, int slow_ld, const float *omega,
                               float kappa0, float dkappa, int nk, const wlsq_c32 *table_forw,
                               wlsq_c32 *forw_pulse, const wlsq_c32 *table_back,
                               wlsq_c32 *back_pulse, float *image, int write_back,
                               const wlsq_tuning *tuning, void *stream)
{
    WLSQ_TRY(check_common(ns, nextrap, nz, nt, nf, nx, M));
    if (ns == 0 || nextrap == 0 || nx == 0)
        return WLSQ_OK;
    if (slow_ld < nx)
        return fail(WLSQ_EINVAL, "slow_ld (%d) < nx (%d)", slow_ld, nx);
    wlsq_tuning t{};
    WLSQ_TRY(workspace_tuning(tuning, &t));
    if (!image || (nf > 0 && (!slowness || !omega || !table_forw || !forw_pulse || !back_pulse ||
                              (!table_back && !t.tables_conjugate))))
        return fail(WLSQ_EINVAL, "NULL device pointer");
    if (nf > 0)
        WLSQ_TRY(check_kappa(nk, dkappa, 2 * M + 1));
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::shared_ptr<PlanImpl> P;
    WsLayout w;
    if (nf > 0) {
        auto shots = uniform_shots(ns, nextrap, nz, nt, nf, nx, slow_ld);
        WLSQ_TRY(build_plan(d, shots.data(), ns, M, &t, P, /*upload=*/false));
        w = ws_layout(*P, nk);
        if (w.need > kNoAllocScratchBytes)
            return fail(WLSQ_ENOMEM,
                        "the allocation-free entry needs %zu bytes of scratch (table copies %zu, "
                        "descriptors %zu), above its maximum of %zu", w.need,
                        (P->shape.conj ? 1 : 2) * w.t_bytes, w.g_bytes + w.i_bytes,
                        kNoAllocScratchBytes);
    }
    unsigned char *base = nullptr;
    WLSQ_TRY(device_scratch(d, &base));
    // image rows l < nextrap are outputs: clear them, then accumulate
    for (int s = 0; s < ns; ++s)
        HIP_TRY(launch_fill_f32(image + (long long)s * nz * nx, (long long)nextrap * nx, 0.f, st));
    if (!P)
        return WLSQ_OK;
    return launch_in_workspace(*P, base, w, slowness, omega, kappa0, dkappa, nk, table_forw,
                               table_back, forw_pulse, back_pulse, image, write_back, st);
}

long long wlsq_batch_workspace_bytes(const wlsq_shot *shots, int nshots, int M, int nk,
                                     const wlsq_tuning *tuning)
{
    wlsq_tuning t{};
    if (workspace_tuning(tuning, &t) != WLSQ_OK)
        return -1;
    Device *d = nullptr;
    if (current_device(&d) != WLSQ_OK)
        return -1;
    std::shared_ptr<PlanImpl> P;
    // the task list does not depend on the steps of a round: size it as a
    // round of every remaining step (the full launch when all first_step are 0)
    if (build_plan(d, shots, nshots, M, &t, P, /*upload=*/false, kMaxRoundSteps) != WLSQ_OK)
        return -1;
    return (long long)ws_layout(*P, nk).need;
}

int wlsq_batch_device_ws(const wlsq_shot *shots, int nshots, int M, const float *slowness,
                         const float *omega, float kappa0, float dkappa, int nk,
                         const wlsq_c32 *table_forw, const wlsq_c32 *table_back, wlsq_c32 *forw,
                         wlsq_c32 *back, float *image, int write_back, const wlsq_tuning *tuning,
                         void *workspace, long long workspace_bytes, void *stream)
{
    wlsq_tuning t{};
    WLSQ_TRY(workspace_tuning(tuning, &t));
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    std::shared_ptr<PlanImpl> P;
    WLSQ_TRY(build_plan(d, shots, nshots, M, &t, P, /*upload=*/false));
    if (P->items.empty())
        return WLSQ_OK;
    if (!slowness || !omega || !table_forw || (!table_back && !P->shape.conj) || !forw || !back ||
        !image || !workspace)
        return fail(WLSQ_EINVAL, "NULL device pointer");
    WLSQ_TRY(check_kappa(nk, dkappa, P->L));
    if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
        return fail(WLSQ_EINVAL, "workspace must be 256-byte aligned");
    const WsLayout w = ws_layout(*P, nk);
    if (workspace_bytes < 0 || (size_t)workspace_bytes < w.need)
        return fail(WLSQ_ENOMEM, "workspace of %lld bytes, the batch needs %zu", workspace_bytes,
                    w.need);
    return launch_in_workspace(*P, static_cast<unsigned char *>(workspace), w, slowness, omega,
                               kappa0, dkappa, nk, table_forw, table_back, forw, back, image,
                               write_back, static_cast<hipStream_t>(stream));
}

int wlsq_batch_round_ws(const wlsq_shot *shots, int nshots, int steps, int M,
                        const float *slowness, const float *omega, float kappa0, float dkappa,
                        int nk, const wlsq_c32 *table_forw, const wlsq_c32 *table_back,
                        wlsq_c32 *forw, wlsq_c32 *back, float *image, const wlsq_tuning *tuning,
                        void *workspace, long long workspace_bytes, void *stream)
{
    if (steps < 1 || steps > kMaxRoundSteps)
        return fail(WLSQ_EINVAL, "steps must be 1..%d (got %d)", kMaxRoundSteps, steps);
    wlsq_tuning t{};
    WLSQ_TRY(workspace_tuning(tuning, &t));
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    std::shared_ptr<PlanImpl> P;
    WLSQ_TRY(build_plan(d, shots, nshots, M, &t, P, /*upload=*/false, steps));
    if (P->items.empty())
        return WLSQ_OK;
    if (!slowness || !omega || !table_forw || (!table_back && !P->shape.conj) || !forw || !back ||
        !image || !workspace)
        return fail(WLSQ_EINVAL, "NULL device pointer");
    WLSQ_TRY(check_kappa(nk, dkappa, P->L));
    if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
        return fail(WLSQ_EINVAL, "workspace must be 256-byte aligned");
    const WsLayout w = ws_layout(*P, nk);
    if (workspace_bytes < 0 || (size_t)workspace_bytes < w.need)
        return fail(WLSQ_ENOMEM, "workspace of %lld bytes, the round needs %zu", workspace_bytes,
                    w.need);
    // the fields carry each shot's state from one round to the next
    return launch_in_workspace(*P, static_cast<unsigned char *>(workspace), w, slowness, omega,
                               kappa0, dkappa, nk, table_forw, table_back, forw, back, image,
                               /*write_back=*/1, static_cast<hipStream_t>(stream));
}

int extrapolate_ktable(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                       const float *slowness, const float *omega, float kappa0, float dkappa,
                       int nk, const wlsq_c32 *table_forw, wlsq_c32 *forw_pulse,
                       const wlsq_c32 *table_back, wlsq_c32 *back_pulse, float *image)
{
    WLSQ_TRY(check_common(ns, nextrap, nz, nt, nf, nx, M));
    if (ns == 0 || nextrap == 0 || nx == 0)
        return WLSQ_OK;
    if (!slowness || !omega || !table_forw || !table_back || !forw_pulse || !back_pulse || !image)
        return fail(WLSQ_EINVAL, "NULL array");
    WLSQ_TRY(check_kappa(nk, dkappa, 2 * M + 1));
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    hipStream_t st = nullptr;
    WLSQ_TRY(host_stream(d, &st));
    const int L = 2 * M + 1;
    const long long plane = (long long)nf * nx;
    std::lock_guard<std::mutex> call_lk(d->host_call_mu);
    Scratch *scp;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        scp = &scratch_for(d, st);
    }
    Scratch &sc = *scp;
    WLSQ_TRY(sc.slow.reserve(sizeof(float) * (size_t)nextrap * nx));
    WLSQ_TRY(sc.omega.reserve(sizeof(float) * (size_t)std::max(nf, 1)));
    WLSQ_TRY(sc.tf.reserve(sizeof(float2) * (size_t)nk * L));
    WLSQ_TRY(sc.tb.reserve(sizeof(float2) * (size_t)nk * L));
    WLSQ_TRY(sc.fields_f.reserve(sizeof(float2) * (size_t)std::max(1ll, ns * plane)));
    WLSQ_TRY(sc.fields_b.reserve(sizeof(float2) * (size_t)std::max(1ll, ns * plane)));
    WLSQ_TRY(sc.img.reserve(sizeof(float) * (size_t)ns * nextrap * nx));
    HIP_TRY(hipMemcpyAsync(sc.slow.p, slowness, sizeof(float) * (size_t)nextrap * nx,
                           hipMemcpyHostToDevice, st));
    if (nf > 0)
        HIP_TRY(hipMemcpyAsync(sc.omega.p, omega, sizeof(float) * (size_t)nf, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(sc.tf.p, table_forw, sizeof(float2) * (size_t)nk * L, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(sc.tb.p, table_back, sizeof(float2) * (size_t)nk * L, hipMemcpyHostToDevice, st));
    const size_t row_bytes = sizeof(float2) * (size_t)plane;
    const size_t shot_pitch = sizeof(float2) * (size_t)nt * nx;
    if (nf > 0) {
        HIP_TRY(hipMemcpy2DAsync(sc.fields_f.p, row_bytes, forw_pulse, shot_pitch, row_bytes, ns,
                                 hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpy2DAsync(sc.fields_b.p, row_bytes, back_pulse, shot_pitch, row_bytes, ns,
                                 hipMemcpyHostToDevice, st));
    }
    wlsq_tuning t{};
    t.tables_conjugate = host_tables_conjugate(table_forw, table_back, (long long)nk * L) ? 1 : 0;
    t.tables_symmetric =
        host_tables_symmetric(table_forw, nk, L) && host_tables_symmetric(table_back, nk, L) ? 1 : 0;
    // device layout of the fields: (ns, nf, nx), i.e. "nt" = nf; image (ns, nextrap, nx)
    WLSQ_TRY(ktable_device_impl(d, ns, nextrap, nextrap, nf, nf, nx, M, sc.slow.as<float>(), nx,
                                sc.omega.as<float>(), kappa0, dkappa, nk,
                                static_cast<const wlsq_c32 *>(sc.tf.p), static_cast<wlsq_c32 *>(sc.fields_f.p),
                                static_cast<const wlsq_c32 *>(sc.tb.p), static_cast<wlsq_c32 *>(sc.fields_b.p),
                                sc.img.as<float>(), 1, &t, st));
    if (nf > 0) {
        HIP_TRY(hipMemcpy2DAsync(forw_pulse, shot_pitch, sc.fields_f.p, row_bytes, row_bytes, ns,
                                 hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpy2DAsync(back_pulse, shot_pitch, sc.fields_b.p, row_bytes, row_bytes, ns,
                                 hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpy2DAsync(image, sizeof(float) * (size_t)nz * nx, sc.img.p,
                             sizeof(float) * (size_t)nextrap * nx, sizeof(float) * (size_t)nextrap * nx,
                             ns, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WLSQ_OK;
}

int wlsq_expand_table_device(int nextrap, int nf, int nx, int M, const float *slowness, int slow_ld,
                             const float *omega, float kappa0, float dkappa, int nk,
                             const wlsq_c32 *table, wlsq_c32 *out, int layout, void *stream)
{
    if (nextrap < 0 || nf < 0 || nx < 0 || M < 0)
        return fail(WLSQ_EINVAL, "negative size");
    if (layout != WLSQ_LAYOUT_NATURAL && layout != WLSQ_LAYOUT_TAPMAJOR)
        return fail(WLSQ_EINVAL, "unknown table layout %d", layout);
    if (nextrap == 0 || nf == 0 || nx == 0)
        return WLSQ_OK;
    if (nf > 65535 || nextrap > 65535)
        return fail(WLSQ_EUNSUPPORTED, "nf and nextrap must be <= 65535");
    if (slow_ld < nx)
        return fail(WLSQ_EINVAL, "invalid slowness stride %d < nx %d", slow_ld, nx);
    WLSQ_TRY(check_kappa(nk, dkappa, 0));
    if (!slowness || !omega || !table || !out)
        return fail(WLSQ_EINVAL, "NULL device pointer");
    HIP_TRY(launch_expand(slowness, slow_ld, omega, kappa0, inv_dkappa(dkappa), nk,
                          reinterpret_cast<const float2 *>(table), reinterpret_cast<float2 *>(out),
                          nextrap, nf, nx, 2 * M + 1, layout, static_cast<hipStream_t>(stream)));
    return WLSQ_OK;
}

int wlsq_probe_fma(int blocks, int iters, float *out, void *stream)
{
    if (!out || blocks <= 0 || iters < 0)
        return fail(WLSQ_EINVAL, "probe: invalid arguments (blocks=%d iters=%d)", blocks, iters);
    HIP_TRY(launch_probe_fma(blocks, iters, out, static_cast<hipStream_t>(stream)));
    return WLSQ_OK;
}

int wlsq_probe_exact_sum(const float *terms, long long n, int m, float *out, void *stream)
{
    if (!out || (n > 0 && !terms) || n < 0 || m < 0 || (m > 0 && n > (1ll << 24)))
        return fail(WLSQ_EINVAL, "exact sum probe: invalid arguments (n=%lld m=%d)", n, m);
    if (m == 0)
        return WLSQ_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned long long *acc = nullptr;
    HIP_TRY(hipMalloc(&acc, sizeof(unsigned long long) * kExactPlanes * (size_t)m));
    hipError_t e = launch_exact_probe(terms, n, m, acc, out, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    (void)hipFree(acc);
    if (e != hipSuccess)
        return fail(WLSQ_EHIP, "exact sum probe: %s", hipGetErrorString(e));
    return WLSQ_OK;
}

long long wlsq_batch_points(const wlsq_shot *shots, int nshots)
{
    long long p = 0;
    for (int i = 0; i < nshots; ++i)
        p += (long long)shots[i].nx * shots[i].nz * shots[i].nf;
    return p;
}

int wlsq_plan_create(wlsq_plan *out, const wlsq_shot *shots, int nshots, int M,
                     const wlsq_tuning *tuning)
{
    if (!out)
        return fail(WLSQ_EINVAL, "NULL plan pointer");
    *out = nullptr;
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    std::shared_ptr<PlanImpl> P;
    WLSQ_TRY(build_plan(d, shots, nshots, M, tuning, P));
    *out = new wlsq_plan_s{P};
    return WLSQ_OK;
}

int wlsq_plan_info(wlsq_plan plan, long long *points, int *n_items, int *block, int *lds_bytes,
                   int *x_per_thread, int *shots_per_group)
{
    if (!plan || !plan->impl)
        return fail(WLSQ_EINVAL, "NULL plan");
    const PlanImpl &P = *plan->impl;
    if (points)
        *points = P.points;
    if (n_items)
        *n_items = (int)P.items.size();
    if (block)
        *block = P.shape.block;
    if (lds_bytes)
        *lds_bytes = (int)P.lds;
    if (x_per_thread)
        *x_per_thread = P.shape.R;
    if (shots_per_group)
        *shots_per_group = P.shape.S;
    return WLSQ_OK;
}

int wlsq_plan_launch(wlsq_plan plan, const float *slowness, const float *omega, float kappa0,
                     float dkappa, int nk, const wlsq_c32 *table_forw, const wlsq_c32 *table_back,
                     wlsq_c32 *forw, wlsq_c32 *back, float *image, int write_back, void *stream)
{
    if (!plan || !plan->impl)
        return fail(WLSQ_EINVAL, "NULL plan");
    return launch_plan(*plan->impl, slowness, omega, kappa0, dkappa, nk, table_forw, table_back,
                       forw, back, image, write_back, static_cast<hipStream_t>(stream));
}

int wlsq_plan_destroy(wlsq_plan plan)
{
    delete plan;
    return WLSQ_OK;
}

int wlsq_batch_device(const wlsq_shot *shots, int nshots, int M, const float *slowness,
                      const float *omega, float kappa0, float dkappa, int nk,
                      const wlsq_c32 *table_forw, const wlsq_c32 *table_back, wlsq_c32 *forw,
                      wlsq_c32 *back, float *image, int write_back, const wlsq_tuning *tuning,
                      void *stream)
{
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    std::shared_ptr<PlanImpl> P;
    WLSQ_TRY(build_plan(d, shots, nshots, M, tuning, P));
    hipStream_t st = static_cast<hipStream_t>(stream);
    WLSQ_TRY(launch_plan(*P, slowness, omega, kappa0, dkappa, nk, table_forw, table_back, forw,
                         back, image, write_back, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WLSQ_OK;
}

int wlsq_graph_create(wlsq_graph *out, wlsq_plan plan, const float *slowness, const float *omega,
                      float kappa0, float dkappa, int nk, const wlsq_c32 *table_forw,
                      const wlsq_c32 *table_back, wlsq_c32 *forw, wlsq_c32 *back, float *image,
                      int write_back, long long zero_count)
{
    if (!out || !plan || !plan->impl)
        return fail(WLSQ_EINVAL, "NULL plan or output");
    *out = nullptr;
    Device *d = nullptr;
    WLSQ_TRY(current_device(&d));
    // no allocation inside the capture: table copies, the field copies of
    // split launches without write-back, and the image planes
    if (nk > 0)
        WLSQ_TRY(plan->impl->reserve_tables(nk));
    {
        PlanImpl &P = *plan->impl;
        if (!write_back && P.depth_steps > 0 && P.depth_steps < P.max_nz)
            WLSQ_TRY(P.fields.reserve(sizeof(float2) * (size_t)(P.f_extent + P.b_extent)));
        if (P.nplanes > 1)
            WLSQ_TRY(P.planes.reserve(sizeof(float) * (size_t)P.nplanes * (size_t)P.img_extent));
        if (P.exact_image)
            WLSQ_TRY(P.xacc.reserve(sizeof(unsigned long long) * kExactPlanes * (size_t)P.img_extent));
    }
    hipStream_t cap = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    auto g = std::make_unique<wlsq_graph_s>();
    g->plan = plan->impl;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    int rc = WLSQ_OK;
    if (e == hipSuccess) {
        if (zero_count > 0) {
            hipError_t fe = launch_fill_f32(image, zero_count, 0.f, cap);
            if (fe != hipSuccess)
                rc = fail(WLSQ_EHIP, "fill in capture: %s", hipGetErrorString(fe));
        }
        if (rc == WLSQ_OK)
            rc = launch_plan(*plan->impl, slowness, omega, kappa0, dkappa, nk, table_forw,
                             table_back, forw, back, image, write_back, cap);
        hipError_t ee = hipStreamEndCapture(cap, &g->graph);
        if (rc == WLSQ_OK && ee != hipSuccess)
            rc = fail(WLSQ_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    } else {
        rc = fail(WLSQ_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    (void)hipStreamDestroy(cap);
    if (rc != WLSQ_OK) {
        if (g->graph)
            (void)hipGraphDestroy(g->graph);
        return rc;
    }
    HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->plan->graphs.fetch_add(1);
    *out = g.release();
    return WLSQ_OK;
}

int wlsq_graph_launch(wlsq_graph graph, void *stream)
{
    if (!graph || !graph->exec)
        return fail(WLSQ_EINVAL, "NULL graph");
    HIP_TRY(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    return WLSQ_OK;
}

int wlsq_graph_destroy(wlsq_graph graph)
{
    if (!graph)
        return WLSQ_OK;
    if (graph->exec)
        (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph)
        (void)hipGraphDestroy(graph->graph);
    if (graph->exec && graph->plan)
        graph->plan->graphs.fetch_sub(1);
    delete graph;
    return WLSQ_OK;
}

}  // extern "C"
