/*
 * libwlsq_hip.so - C ABI of the MI355X (gfx950) extrapolate-and-image library.
 *
 * Every entry point returns an int status (WLSQ_OK = 0). The reference's
 * entries return void and ctypes binds them with restype = c_void_p
 * (cpp_src/interface_cuda.py:31), so an int status is compatible with that
 * binding; wlsq_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Types: wlsq_c32 is complex64 (8 bytes, interleaved re, im), the layout of
 * thrust::complex<float> in cpp_src/types.h:4-9 and of numpy.complex64.
 * Sizes are int as in the reference; all offsets inside are 64-bit.
 *
 * "host" entries take host arrays and copy in and out (the reference's
 * contract: caller-owned numpy buffers, cpp_src/extrapolation_revOp_gpu.cu:78-231);
 * "_device" entries take device pointers and a hipStream_t passed as void*
 * and return once the work is queued on that stream.
 */
#ifndef WLSQ_HIP_H
#define WLSQ_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float re, im;
} wlsq_c32;

enum {
    WLSQ_OK = 0,
    WLSQ_EINVAL = 1,      /* invalid argument (shape, pointer, layout) */
    WLSQ_EHIP = 2,        /* a HIP runtime call failed */
    WLSQ_ENOMEM = 3,      /* device allocation failed or over the memory budget */
    WLSQ_EUNSUPPORTED = 4 /* shape outside what the kernels support (e.g. LDS) */
};

/* operator-table layouts of the reference-shaped ("expanded") tables */
enum {
    WLSQ_LAYOUT_NATURAL = 0, /* (nextrap, nf, nx, L): preparation.py:103, the caller's layout */
    WLSQ_LAYOUT_TAPMAJOR = 1 /* (nextrap, nf, L, nx): cpp_src/revOp.c:4-30 ("revOp") */
};

const char *wlsq_last_error(void);
int wlsq_version(void);

/* Device ordinal used by every later call of this thread (default 0). */
int wlsq_set_device(int device);

/*
 * Drop-in replacement of the GPU entry
 *   void extrapolate(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
 *                    fcomp* w_op_forw, fcomp* forw_pulse, fcomp* w_op_back,
 *                    fcomp* back_pulse, float* image)
 * of cpp_src/extrapolation_revOp_gpu.cu:78-80 (same signature in
 * cpp_src/extrapolation_gpu.cu:74-76 and cpp_src/extrapolation_sh_revOp_gpu.cu:87-89),
 * bound by cpp_src/interface_cuda.py:30-38 and called from main.py:138-150.
 *   w_op_forw/back: (nextrap, nf, nx, 2M+1) host, natural layout
 *   forw/back_pulse: (ns, nt, nx) host; rows j < nf are read and overwritten
 *                    with the final wavefields (shot stride nt*nx, as
 *                    cpp_src/extrapolation_cpu.cpp:91-98)
 *   image: (ns, nz, nx) host; rows l < nextrap are written with
 *          Re(sum_j F conj(B)) after step l; rows >= nextrap are untouched.
 * Synchronous, like the reference (cudaDeviceSynchronize at :180).
 * interface_cuda.py binds the symbol `extrapolate` of three libraries under
 * three Python names (:16-43); this one symbol serves all three.
 */
int extrapolate(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                wlsq_c32 *w_op_forw, wlsq_c32 *forw_pulse, wlsq_c32 *w_op_back,
                wlsq_c32 *back_pulse, float *image);

/*
 * The plain path on device pointers ("one launch sequence per shot" when
 * shots_per_launch == 1): per depth step one fused kernel updates both fields
 * of shots_per_launch shots for every (omega, x), streaming the expanded
 * tables; per-depth imaging partials are reduced in fixed order at the end.
 * Replaces the hot loop of cpp_src/extrapolation_revOp_gpu.cu:142-179
 * (kernels :16-72). Tables in `layout` (natural tables are transposed on the
 * device into a library scratch buffer first, cpp_src/revOp.c:4-30).
 * shots_per_launch: 1, 2 or 4 (shots sharing one coefficient read).
 */
int wlsq_plain_device(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                      const wlsq_c32 *op_forw, wlsq_c32 *forw_pulse,
                      const wlsq_c32 *op_back, wlsq_c32 *back_pulse, float *image,
                      int layout, int shots_per_launch, void *stream);

/*
 * Compact-table ("k-table") form of the same computation: the per-point
 * operator is gathered on the device from (nk, 2M+1) tables by
 *   q = nearest kappa sample to omega_j * slowness[l][i]
 * (the lookup of py_src/operatorsPreparation/preparation.py:109 with
 * util.py:5-8; kappa_q = kappa0 + q*dkappa; ties -> lower q; evaluated in
 * FP32 as t = omega*s; t -= kappa0; u = t*inv_dk; clamp [0, nk-1];
 * q = ceil(u - 0.5) with inv_dk = (float)(1.0/(double)dkappa)).
 * Gives bit-identical tables to expanding with the same rule
 * (wlsq_expand_table_device), i.e. the same images as `extrapolate` on the
 * expanded table.
 *   slowness: (>= nextrap rows, slow_ld) f32, row l used at step l, columns 0..nx-1
 *   omega: (nf) f32 rad/s
 * The fused kernel keeps each (shot group, omega group) row resident in LDS
 * for all nextrap steps and accumulates the image with atomics.
 */
int extrapolate_ktable(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                       const float *slowness, const float *omega, float kappa0, float dkappa,
                       int nk, const wlsq_c32 *table_forw, wlsq_c32 *forw_pulse,
                       const wlsq_c32 *table_back, wlsq_c32 *back_pulse, float *image);

typedef struct {
    int shots_per_group; /* 0 = auto, else 1, 2 or 4 shots sharing coefficient loads */
    int omegas_per_item; /* 0 = auto: frequencies per workgroup task, at most
                            what the threads and the LDS budget allow; with
                            1024 threads each shot's frequencies are split
                            into tasks that balance the four SIMDs (so the
                            tasks of one shot may differ by a few rows);
                            n > 0 = at most n, and every task of a shot but
                            the last takes the same count */
    int block;           /* 0 = auto: threads per workgroup, a multiple of 64:
                            up to 1024 for one shot per group at 4 x per
                            thread and for two shots at 2 x, else up to 512 */
    int x_per_thread;    /* 0 = auto: consecutive x per thread, 2 (with 2 or 4
                            shots per group), 4 or 8 (8 with 1 shot per
                            group); a task's aperture is covered by one or two
                            strips of x_per_thread per thread (two: 4 x at
                            more than 512 threads, or 8 x), up to 8192 */
    int lds_budget;      /* 0 = auto: bytes of LDS per workgroup */
    int tables_conjugate; /* 1: caller guarantees table_back == conj(table_forw)
                             bit for bit, so only table_forw is read */
    int image_mode;       /* 0 = auto (4 with one shot per group; else 1 where
                             the partial rows fit next to the field rows,
                             elsewhere 3), 1 = per-frequency image
                             partial rows in LDS, summed in fixed order, then one
                             atomic add per (depth, x) per task, 2 = an atomic add
                             per point from the thread that computed it, 3 = as 2,
                             but with one shot per group, 4 x per thread and
                             more than 512 threads, single-frequency tasks first
                             transpose each wavefront's values through a
                             wave-private LDS segment (64 x 4 floats per
                             wavefront), so each atomic instruction covers
                             consecutive x, and the tasks add into image planes
                             (image_planes); other shapes add as in 2;
                             4 = as 1, but the sum over the task's
                             frequencies (and shots sharing an image) is
                             formed from the new field rows after each step,
                             in the same order, without partial rows: more
                             frequency rows fit a task (one shot per group) */
    int tables_symmetric; /* 1: caller guarantees both tables are tap-symmetric,
                             T[q][k] == T[q][2M-k] bit for bit (the WLSQ design
                             with symmetric taps gives this); taps are then
                             paired: (old[x+k-M] + old[x+M-k]) * T[k], which
                             changes the rounding order, not the operator */
    int tap_loads;        /* 0 = auto (2 with one shot per group and up to 4 x
                             per thread, else 1), 1 = one tap per coefficient
                             load, 2 = two taps per 16-byte load */
    int strip_gathers;    /* 0 = auto (on), 1 = on: where every lane of a
                             wavefront sees at most one change of kappa row
                             over its x_per_thread positions, gather each tap
                             once or twice per lane instead of once per x;
                             2 = off. Results are bitwise the same. */
    int double_rows;      /* 0 = auto (on when the tasks' doubled LDS fits the
                             device), 1 = on (error if it does not fit), 2 =
                             off: two copies of the rows in LDS, one barrier
                             per depth step instead of two */
    int depth_steps;      /* 0 = every depth step in one launch; n > 0 = launches
                             of n steps, each continuing from the fields the
                             previous one wrote (without write-back they run
                             on device copies of the fields: the caller's stay
                             untouched) */
    int image_planes;     /* direct atomics of image_mode 3 (also auto where
                             the partial rows do not fit): 0 = auto (8 planes
                             while they take at most 512 MiB, halved until they
                             fit, none below 2), 1 = none, n <= 16 = n planes.
                             Consecutive tasks add into different image-shaped
                             planes of plan-owned device memory, zeroed on the
                             launch stream and summed into the image after the
                             launch, so tasks running side by side do not
                             contend for the same image addresses. With
                             image_mode 4, n >= 2 planes are used likewise
                             (only when asked for). */
    int rounding;         /* 0 = fast: complex products formed with fused
                             multiply-adds, taps paired where the tables are
                             symmetric; 1 = the CPU path's rounding: every
                             complex product formed with separate multiplies
                             and adds, taps summed in order k = 0..2M from zero
                             (cpp_src/extrapolation_cpu.cpp:52-53 with
                             cpp_src/mkl_linalg.cpp:10), imaging term
                             re*re + im*im. The final fields are then bit for
                             bit those of the CPU path on the same tables;
                             about half the speed. Needs 4 x per thread, one
                             shot per group and more than 512 threads (the
                             planner picks them when those fields are 0);
                             tables_symmetric is ignored. */
    int strips_per_thread; /* 0 = auto (two only where the aperture has more
                              strips than the workgroup has threads), 1, or 2:
                              each thread steps two strips of its task's rows
                              (4 x per thread with one shot per group and more
                              than 512 threads, or 8 x per thread), so a task
                              holds up to twice the frequency rows, in one copy
                              of the rows (two barriers per step). With
                              lds_budget, the autotuner uses it to trade rows
                              per task against tasks. */
    int exact_image;      /* 1: each task's image sums (image modes 1 and 4)
                             are added exactly, as 64-bit integers on the
                             2^-149 grid into 16 words per image point (plus
                             a flags word for infinities and NaN), and rounded
                             once after the launch to the nearest float (ties
                             to even): the image does not depend on the order
                             of the atomic adds, so launches of one plan give
                             it bit for bit, and it is the correctly rounded
                             sum of the task sums. 136 bytes of plan memory
                             per image point; plan launches only (the
                             workspace and allocation-free entries refuse it),
                             image modes 1 and 4, tables_conjugate = 1;
                             2^24 task sums per point at most. */
} wlsq_tuning;

int wlsq_ktable_device(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                       const float *slowness, int slow_ld, const float *omega,
                       float kappa0, float dkappa, int nk,
                       const wlsq_c32 *table_forw, wlsq_c32 *forw_pulse,
                       const wlsq_c32 *table_back, wlsq_c32 *back_pulse, float *image,
                       int write_back, const wlsq_tuning *tuning, void *stream);

/*
 * The computation of wlsq_ktable_device without any device allocation. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture - the launches it queues are all
 * kernels, descriptors travel in kernel arguments -, tight serving loops,
 * harnesses that forbid hipMalloc). The plan's group and task descriptors and
 * the row-padded table copies live in a library-owned device array of
 * 16 MiB per GPU, held until the process ends: the call needs
 * (2M+2)*nk*8 bytes per table read (one when tuning->tables_conjugate) plus
 * 144 bytes per shot group and 16 per task, and returns WLSQ_ENOMEM with the
 * sizes above that maximum. tuning->depth_steps must be 0; image planes are
 * not used (direct image atomics add into `image`). Every call on one GPU
 * uses that GPU's one array: calls that overlap on one GPU (other streams,
 * other threads) race and give wrong results; calls queued on one stream
 * run in order. Same arguments and results as wlsq_ktable_device.
 */
int wlsq_ktable_device_noalloc(int ns, int nextrap, int nz, int nt, int nf, int nx, int M,
                               const float *slowness, int slow_ld, const float *omega,
                               float kappa0, float dkappa, int nk, const wlsq_c32 *table_forw,
                               wlsq_c32 *forw_pulse, const wlsq_c32 *table_back,
                               wlsq_c32 *back_pulse, float *image, int write_back,
                               const wlsq_tuning *tuning, void *stream);

/*
 * Expansion of a compact table into the reference-shaped table
 * (replaces makeTableOfOperators, preparation.py:101-115, for the symmetric
 * operator of every x, main.py:66): out[l][j][i][:] = table[q(l,j,i)][:] in
 * `layout`. slowness row stride slow_ld.
 */
int wlsq_expand_table_device(int nextrap, int nf, int nx, int M, const float *slowness,
                             int slow_ld, const float *omega, float kappa0, float dkappa,
                             int nk, const wlsq_c32 *table, wlsq_c32 *out, int layout,
                             void *stream);

/*
 * Batched form: many shots of uneven aperture, depth count and frequency
 * list in ONE launch (no counterpart in the reference, whose GPU entries take
 * one common shape for all shots, cpp_src/extrapolation_revOp_gpu.cu:78-80).
 * Each shot reads slowness rows from slowness + slow_off (window origin,
 * row stride slow_ld), its frequencies from omega + omega_off, its fields
 * from forw/back + f_off/b_off (row j at + j*ld), and adds its image into
 * image + img_off (row l at + l*img_ld): point every shot at the same image
 * for a stacked image.
 */
typedef struct {
    long long f_off, b_off; /* element offsets of row 0 of the forward/backward fields */
    long long img_off;      /* element offset of image row 0 */
    long long slow_off;     /* element offset of (depth 0, window column 0) in slowness */
    int ld;                 /* field row stride (elements) */
    int img_ld;             /* image row stride */
    int slow_ld;            /* slowness row stride */
    int nx;                 /* aperture (window width) */
    int nz;                 /* depth steps */
    int nf;                 /* frequencies */
    int omega_off;          /* offset of the frequency list in omega */
    int first_step;         /* wlsq_batch_round_ws: the shot's next depth step (its
                               fields hold the wavefields after step first_step - 1);
                               every other entry: must be 0 */
} wlsq_shot;

/* One call: plan, launch, wait, free (synchronous). Adds into `image`. */
int wlsq_batch_device(const wlsq_shot *shots, int nshots, int M,
                      const float *slowness, const float *omega, float kappa0, float dkappa,
                      int nk, const wlsq_c32 *table_forw, const wlsq_c32 *table_back,
                      wlsq_c32 *forw, wlsq_c32 *back, float *image, int write_back,
                      const wlsq_tuning *tuning, void *stream);

/* The batched form without device allocation and without waiting: the work
 * list is built on the host, its descriptors and the row-padded table copies
 * are written into the caller's `workspace` (256-byte aligned, at least
 * wlsq_batch_workspace_bytes) by kernels queued on `stream`, then the fused
 * kernel; returns once queued. For loops that launch many batches (the
 * serving loop keeps one workspace per stream); no counterpart in the
 * reference. depth_steps must be 0; image planes are not used. A workspace
 * must not be reused before the launch that wrote it has finished (queue its
 * next use on the same stream). */
long long wlsq_batch_workspace_bytes(const wlsq_shot *shots, int nshots, int M, int nk,
                                     const wlsq_tuning *tuning); /* -1 on invalid input */
int wlsq_batch_device_ws(const wlsq_shot *shots, int nshots, int M, const float *slowness,
                         const float *omega, float kappa0, float dkappa, int nk,
                         const wlsq_c32 *table_forw, const wlsq_c32 *table_back, wlsq_c32 *forw,
                         wlsq_c32 *back, float *image, int write_back, const wlsq_tuning *tuning,
                         void *workspace, long long workspace_bytes, void *stream);

/* One round of continuous batching (no counterpart in the reference; for
 * serving loops that let new jobs join running work): each shot runs depth
 * steps [first_step, min(nz, first_step + steps)) from the wavefields its
 * rows hold, writes them back (the state the next round continues from) and
 * adds image rows first_step.. into `image`; shots with no step left have no
 * work. Shots of different jobs (own fields, image rows, frequency lists and
 * first steps) share the launch, so jobs admitted between rounds run beside
 * the ones already running. Workspace, tuning and stream rules as for
 * wlsq_batch_device_ws (wlsq_batch_workspace_bytes sizes it for any round of
 * these shots); steps 1..2^24. */
int wlsq_batch_round_ws(const wlsq_shot *shots, int nshots, int steps, int M,
                        const float *slowness, const float *omega, float kappa0, float dkappa,
                        int nk, const wlsq_c32 *table_forw, const wlsq_c32 *table_back,
                        wlsq_c32 *forw, wlsq_c32 *back, float *image, const wlsq_tuning *tuning,
                        void *workspace, long long workspace_bytes, void *stream);

/* Reusable plan: the work list (shot groups x frequency groups, longest
 * first) and launch shape are built once and kept on the device. */
typedef struct wlsq_plan_s *wlsq_plan;
int wlsq_plan_create(wlsq_plan *out, const wlsq_shot *shots, int nshots, int M,
                     const wlsq_tuning *tuning);
/* points, workgroup tasks, threads per workgroup, LDS bytes, x per thread,
 * shots per group; any pointer may be NULL */
int wlsq_plan_info(wlsq_plan plan, long long *points, int *n_items, int *block, int *lds_bytes,
                   int *x_per_thread, int *shots_per_group);
/* Queue the plan's launch on `stream` (a row-padded copy of the tables into
 * the plan's storage, then the fused kernel, or one kernel per depth range
 * with tuning.depth_steps); adds into `image`. Launches of one plan on
 * different streams must not overlap: the table copies, and the field copies
 * of split launches without write-back, belong to the plan. */
int wlsq_plan_launch(wlsq_plan plan, const float *slowness, const float *omega, float kappa0,
                     float dkappa, int nk, const wlsq_c32 *table_forw,
                     const wlsq_c32 *table_back, wlsq_c32 *forw, wlsq_c32 *back, float *image,
                     int write_back, void *stream);
int wlsq_plan_destroy(wlsq_plan plan);

/* hipGraph of [zero `zero_count` floats at `image` (fill kernel, no memset
 * node)] + the plan's launch, captured once and replayed by
 * wlsq_graph_launch. The pointers are baked into the graph. */
typedef struct wlsq_graph_s *wlsq_graph;
int wlsq_graph_create(wlsq_graph *out, wlsq_plan plan, const float *slowness, const float *omega,
                      float kappa0, float dkappa, int nk, const wlsq_c32 *table_forw,
                      const wlsq_c32 *table_back, wlsq_c32 *forw, wlsq_c32 *back, float *image,
                      int write_back, long long zero_count);
int wlsq_graph_launch(wlsq_graph graph, void *stream);
int wlsq_graph_destroy(wlsq_graph graph);

/* Measurement aid, no counterpart in the reference: `blocks` workgroups of
 * 256 threads (4 waves) whose every wave issues exactly 8 * iters FP32 FMA
 * instructions in eight independent chains (plus a 7-add tail and one store
 * of a float per thread into out[blocks * 256]). Its known instruction count
 * calibrates the VALU counters of a profile, and its time gives the FMA
 * issue rate the device reaches with packed FP32 disabled. */
int wlsq_probe_fma(int blocks, int iters, float *out, void *stream);

/* Measurement aid, no counterpart in the reference: out[j] = the float32
 * nearest (ties to even) the exact sum over i < n of terms[i * m + j], the
 * terms added by many threads in no fixed order through the exact image
 * accumulation of the fused kernel (exact_image); device pointers, n <= 2^24;
 * synchronous (allocates its accumulators). */
int wlsq_probe_exact_sum(const float *terms, long long n, int m, float *out, void *stream);

/* Points (x, z, omega, shot) the batched form computes for these shots. */
long long wlsq_batch_points(const wlsq_shot *shots, int nshots);

/* ---------------------------------------------------------------- transforms
 * In-place complex FFT of length N of a host array, computed on the device
 * (hipFFT). Drop-ins for the reference CPU library's MKL wrappers
 * c64fft1dforw / c64fft1dback (cpp_src/mkl_fft.cpp:9-39, bound in
 * cpp_src/interface.py:36-45): forward X[k] = sum_n x[n] exp(-2 pi i k n / N);
 * backward scaled by 1/N. Returns a status instead of void. */
int c64fft1dforw(wlsq_c32 *data, int N);
int c64fft1dback(wlsq_c32 *data, int N);

/* Spectra of real traces on the device: out[j * ld_out + x] =
 * sum_t traces[t * ld_in + x] exp(-2 pi i j t / nt) for all j < nt, x < nx,
 * i.e. scipy.fftpack.fft(traces, axis=0) of the shot and source records in
 * main.py:119-123, without the host FFT and the upload of complex spectra.
 * Rows above nt/2 come from conjugate symmetry. `out` must not overlap
 * `traces`. Enqueued on `stream` (0 = null stream); plans are cached. */
int wlsq_trace_spectra_device(int nt, int nx, const float *traces, int ld_in, wlsq_c32 *out,
                              int ld_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WLSQ_HIP_H */
