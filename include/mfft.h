/*
 * mfft -- MI355X (gfx950) matrix-core FFT and FFT-convolution library, C ABI.
 *
 * This is the native core under the reference's plan/exec API (tcfft_half.h,
 * tcfft_half_2d.h, tensorcore_fft_conv2d.h; see those headers in this directory for the
 * drop-in names). All entry points are extern "C", take plain pointers and sizes and
 * return an int status (MFFT_OK = 0, negative on error; mfft_last_error() gives the
 * message of the calling thread's last failure). Device pointers are HIP device memory;
 * streams are hipStream_t passed as void* (NULL = the default stream).
 *
 * Conventions (reference: main.cpp:68-94, tensorcore_fft_conv2d_kernel.cu:47-82):
 *   - complex data is interleaved (re, im), element index fastest, then batch;
 *   - MFFT_FORWARD computes sum_n x[n] exp(-2*pi*i*n*k/N), MFFT_INVERSE uses +i;
 *   - 2D data is [batch][nx][ny] with y contiguous (tcfft_half_2d.cu:1264-1271);
 *   - the output equals out_scale * DFT(in), where out_scale (mfft_plan_info) is 1 for
 *     MFFT_SCALE_NONE (the reference's unnormalised contract) and 1/(product of the
 *     transform lengths) for MFFT_SCALE_STAGE (every radix-r stage multiplies by 1/r,
 *     so no intermediate can exceed the input's magnitude -- the fp16 overflow guard).
 */
#ifndef MFFT_H
#define MFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MFFT_OK = 0,
    MFFT_ERR_INVALID = -1,     /* bad argument */
    MFFT_ERR_UNSUPPORTED = -2, /* size/format combination not provided */
    MFFT_ERR_HIP = -3,         /* a HIP runtime call failed */
    MFFT_ERR_NOMEM = -4,       /* device allocation failed or budget exceeded */
};

/* arithmetic of the radix-16 stages */
enum {
    MFFT_PREC_FP16 = 0,       /* fp16 matrix-core operands, fp32 accumulate */
    MFFT_PREC_BF16 = 1,       /* bf16 matrix-core operands, fp32 accumulate */
    MFFT_PREC_SPLIT_FP16 = 2, /* split fp16 hi+lo operands, 3 products, fp32 accumulate */
    MFFT_PREC_FP32 = 3,       /* plain fp32 vector-ALU butterflies (comparator) */
};

/* complex element formats */
enum {
    MFFT_FMT_C16 = 0,   /* 2 x fp16 (the reference's half2) */
    MFFT_FMT_CBF16 = 1, /* 2 x bf16 */
    MFFT_FMT_C32 = 2,   /* 2 x fp32 */
};

enum { MFFT_FORWARD = -1, MFFT_INVERSE = 1 };
enum { MFFT_SCALE_NONE = 0, MFFT_SCALE_STAGE = 1 };
/* input order: natural, or the reference's mixed-radix digit-reversed order
 * (tcfft_half.cu:1351-1374, applied by setup() at tcfft_half.cu:1385-1391) */
enum { MFFT_ORDER_NATURAL = 0, MFFT_ORDER_DIGIT_REVERSED = 1 };

typedef struct mfft_desc {
    int precision;            /* MFFT_PREC_* */
    int direction;            /* MFFT_FORWARD / MFFT_INVERSE */
    int scale;                /* MFFT_SCALE_* */
    int in_fmt, out_fmt;      /* MFFT_FMT_* */
    int in_order;             /* MFFT_ORDER_* */
    int64_t workspace_limit;  /* bytes of intermediate buffer per plan (0 = default) */
    int split_hint;           /* first-pass length of a four-step split (1D; the column axis
                                 of a 2D plan, which is otherwise one pass up to 1024 points):
                                 up to 256, or a power of two up to 1024 (0 = planner's choice;
                                 1 = odd * 2^a lengths in one mixed-radix pass;
                                 MFFT_SPLIT_TUNED = split and workspace_limit from the
                                 autotuner's cache, planner's choice when absent) */
    int mid_fmt;              /* format of intermediates between passes: -1 = the precision's
                                 default (fp16 -> C16, bf16 -> CBF16, split/fp32 -> C32), or an
                                 MFFT_FMT_*; split fp16 with C16 rounds once per pass boundary */
    int external_workspace;   /* 1: the plan allocates no workspace; every exec passes one
                                 (mfft_exec_ex) of at least mfft_info.workspace_bytes */
} mfft_desc;

typedef struct mfft_plan_s *mfft_plan;

#define MFFT_MAX_PASSES 6
typedef struct mfft_pass_info {
    int R;          /* points per line of the pass */
    int lines_per_tile;
    int lane_lines; /* lines interleaved in one matrix-core column group */
    int nstages;
    int radix[8];
    int64_t tiles;  /* workgroups per launch (per batch chunk) */
    int generic;    /* 1 = runtime-schedule kernel (odd radices) */
    int in_elem_bytes, out_elem_bytes; /* bytes per complex element read / written */
    int64_t points; /* complex elements read (= written) per launch of a full chunk */
    int kernel;     /* MFFT_KERNEL_*: which kernel family runs the pass */
} mfft_pass_info;

enum {
    MFFT_KERNEL_TILE = 0,      /* LDS tile kernel, power-of-two schedule (fft_tile_pow2) */
    MFFT_KERNEL_GENERIC = 1,   /* LDS tile kernel, runtime mixed-radix schedule */
    MFFT_KERNEL_ROWS = 2,      /* matrix-core kernel, contiguous lines in and out (fft_fast) */
    MFFT_KERNEL_ROWS_TR = 3,   /* matrix-core kernel, contiguous in, transposed out (fft_fast) */
    MFFT_KERNEL_COLS = 4,      /* matrix-core kernel, strided 256-point lines (fft_fast) */
    MFFT_KERNEL_STREAM = 5,    /* streaming short-DFT kernel, one thread per line (small_cols) */
    MFFT_KERNEL_ONCHIP = 6,    /* one workgroup per 65536-point line, the intermediate on chip (fft_onchip) */
    MFFT_KERNEL_ONCHIP_COLS = 7 /* 32 strided 2048-point columns per workgroup, on chip (fft_onchip_cols) */
};

typedef struct mfft_info {
    int kind;              /* 1 = 1D, 2 = 2D */
    int n, nx, ny;
    int64_t batch;
    int npasses;
    mfft_pass_info pass[MFFT_MAX_PASSES];
    double out_scale;      /* out = out_scale * DFT(in) */
    int64_t workspace_bytes;
    int64_t batch_chunk;   /* transforms per chunk (intermediates stay cache-resident) */
    int64_t launches_per_exec;
} mfft_info;

/* Fill a descriptor with defaults: split fp16, forward, per-stage scaling, fp16 in/out,
 * natural order. */
int mfft_desc_default(mfft_desc *d);

/* Plans own their device tables and workspace; they are reentrant (exec keeps no mutable
 * state besides the workspace, so two streams need two plans). Replaces tcfftCreate
 * (tcfft_half.cu:1235-1345) and tcfftCreate_2d (tcfft_half_2d.cu:1151-1221). */
int mfft_plan_create_1d(mfft_plan *plan, int n, int64_t batch, const mfft_desc *d);
int mfft_plan_create_2d(mfft_plan *plan, int nx, int ny, int64_t batch, const mfft_desc *d);
int mfft_plan_destroy(mfft_plan plan);
int mfft_plan_info(mfft_plan plan, mfft_info *info);

/* Execute on `stream`. in == out is allowed (in place, as tcfftExec). Replaces
 * tcfftExec (tcfft_half.cu:1177-1233), tcfftExec_2d (tcfft_half_2d.cu:1128-1149) and
 * TensorCoreFFTExec (tensorcore_fft_conv2d_kernel.cu:1495-1525). */
int mfft_exec(mfft_plan plan, const void *in, void *out, void *stream);

/* As mfft_exec, and records hipEvent_t events[k] on the stream before launch k and
 * events[launches] after the last one (nevents >= launches_per_exec + 1), so a caller
 * can time every kernel of the exec inside its own timed region. */
int mfft_exec_events(mfft_plan plan, const void *in, void *out, void *stream, void **events, int nevents);

/* Runs pass `pass` (0 .. npasses-1) of an exec on its own, every chunk, on the buffers it
 * uses in a full exec (in, out or the plan's workspace): a pass reads what the previous
 * pass left. start_event / stop_event (hipEvent_t, may be NULL): recorded by the runtime
 * around the pass's kernel itself (hipExtLaunchKernel), i.e. its execution window without
 * the launch gap; needs a plan that runs its batch in one chunk. For per-kernel timing.
 * No counterpart in the reference. */
int mfft_exec_pass(mfft_plan plan, int pass, const void *in, void *out, void *stream, void *start_event,
                   void *stop_event);

/* Work-queue batching: many transforms of uneven length (n <= 4096, radices
 * {16, 8, 4, 2, 3, 5, 17}) and batch count in ONE launch of a persistent kernel that takes
 * tiles from a device-side counter. Items are contiguous batches in the batcher's formats. */
typedef struct mfft_batch_item {
    const void *in;
    void *out;
    int n;
    int64_t count;
} mfft_batch_item;
typedef struct mfft_batcher_s *mfft_batcher;
int mfft_batcher_create(mfft_batcher *b, const mfft_desc *d, int max_items);
int mfft_batcher_exec(mfft_batcher b, const mfft_batch_item *items, int nitems, void *stream);
int mfft_batcher_destroy(mfft_batcher b);

/* As mfft_exec, for the first `batch` transforms only (batch <= the plan's), using the
 * caller's workspace (>= mfft_info.workspace_bytes bytes, or NULL for the plan's own):
 * one plan can then serve several streams at once, each with its own workspace. */
int mfft_exec_ex(mfft_plan plan, const void *in, void *out, int64_t batch, void *workspace, void *stream);

/* As mfft_exec_ex with no workspace argument and no allocation: the plan's intermediates go
 * to a device array the library holds for this entry point, MFFT_NOALLOC_SCRATCH_BYTES per
 * GPU, reserved when the library's code is loaded on the GPU and kept until the process
 * ends. No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The plan is created beforehand (creation allocates its tables); a plan whose
 * mfft_info.workspace_bytes exceeds the array is refused with MFFT_ERR_NOMEM (create it
 * with workspace_limit <= MFFT_NOALLOC_SCRATCH_BYTES). All calls on one GPU share its array:
 * calls that overlap on one GPU race, so callers order them (one stream, or events). */
#define MFFT_NOALLOC_SCRATCH_BYTES (64LL << 20)
int mfft_exec_noalloc(mfft_plan plan, const void *in, void *out, int64_t batch, void *stream);

/* hipGraph-captured exec: the plan's launches for fixed (in, out) pointers are captured
 * once and replayed with one graph launch (no per-launch host work). The graph holds no
 * memset nodes. Destroy before the plan. */
typedef struct mfft_graph_s *mfft_graph;
int mfft_graph_create(mfft_graph *graph, mfft_plan plan, const void *in, void *out, void *stream);
int mfft_graph_launch(mfft_graph graph, void *stream);
int mfft_graph_destroy(mfft_graph graph);

/* Digit-reversal table of the plan's radix list (gen_rev, tcfft_half.cu:1351-1374):
 * position i of a digit-reversed buffer holds natural element rev[i]. For 2D plans,
 * axis 0 = x (rev_x), axis 1 = y (rev_y). */
int mfft_input_permutation(mfft_plan plan, int axis, int *rev, int n);

/* On-device staging (replaces the reference's host-side setup()/finalize() loops,
 * tcfft_half.cu:1385-1405, and its rev / half2double kernels,
 * tensorcore_fft_conv2d_kernel.cu:1277-1285, 1407-1414): converts `batch` blocks of
 * n_inner complex elements from src_fmt to dst_fmt (MFFT_FMT_* or MFFT_FMT_C64 for
 * interleaved fp64), multiplying by `scale`, and gathers dst[b][i] = src[b][perm_x[i]]
 * (1D) or dst[b][x][y] = src[b][perm_x[x]][perm_y[y]] (2D, n_inner = nx * ny) when the
 * device tables are given (NULL = identity). FP64 values are rounded once. */
#define MFFT_FMT_C64 3
int mfft_stage_convert(const void *src, int src_fmt, void *dst, int dst_fmt, int64_t n_inner, int64_t batch,
                       const int *perm_x, int nx, const int *perm_y, int ny, double scale, void *stream);

/* ------------------------------------------------------------------------------------
 * FFT-based 2D convolution operator (replaces class fft_conv2d_half,
 * tensorcore_fft_conv2d.h:41-197, forward path tensorcore_fft_conv2d.cpp:12-58, 250-363).
 * out[b][f] = sum_c in[b][c] (*) filt[f][c] for a batch of B images with C channels and
 * F filters, via zero-padded 2D FFTs: the filter spectra are computed once
 * (mfft_conv2d_set_filters) and stay resident; two real images share one complex
 * transform (real and imaginary parts).
 *   MFFT_CONV_SAME        'same' convolution: output h x w, centred on the full
 *                         convolution at offset ((fh-1)/2, (fw-1)/2), zero boundary
 *   MFFT_CONV_VALID_CORR  the reference's 'valid' cross-correlation: output
 *                         (h-fh+1) x (w-fw+1) (main.cpp:97-116)
 * ------------------------------------------------------------------------------------ */
enum { MFFT_CONV_SAME = 0, MFFT_CONV_VALID_CORR = 1 };
enum { MFFT_REAL_F32 = 0, MFFT_REAL_F64 = 1 };
enum { MFFT_CONV_SPEC_FP32 = 0, MFFT_CONV_SPEC_FP16 = 1, MFFT_CONV_SPEC_FP32_MID_FP16 = 2, MFFT_CONV_SPEC_REFERENCE = 3 };

typedef struct mfft_conv2d_desc {
    int batch, channels, filters;
    int h, w, fh, fw;
    int mode;           /* MFFT_CONV_* */
    int pad_h, pad_w;   /* transform size; 0 = smallest supported size >= h + fh - 1 */
    int precision;      /* MFFT_PREC_* of the transforms (default split fp16) */
    int spec_fp16;      /* MFFT_CONV_SPEC_*. FP32 (0, default): image spectra and the
                           transforms' intermediates in fp32. FP16 (1): both in fp16 (half the
                           bytes of every pass; each pass boundary rounds to fp16, rel-L2 of
                           the outputs about 1e-4..5e-4). FP32_MID_FP16 (2): fp32 image spectra
                           and transform outputs, fp16 words at the pass boundaries inside each
                           transform; single-channel operators, one filter per call (filters
                           = 1 or mfft_conv2d_forward_select). REFERENCE (3): the reference's
                           method (tensorcore_fft_conv2d_kernel.cu:1033-1150): fp16 transforms
                           without scaling, fp16 image spectra, the channel sum in fp32, each
                           output map quantised to fp16 with a power-of-ten scale (1 if its
                           largest |re|, |im| is at most 1000, else 10^(trunc(log10 max) - 1))
                           before the fp16 inverse, dequantised with the 1/(Ph*Pw) readout; the
                           precision field is ignored. Like the reference, the unscaled fp16
                           forward overflows once an image's padded sum exceeds 65504 (256^2
                           images of [0,1] values fit). Only FP32 serves mfft_conv2d_backward, and
                           forward_two takes FP32, FP16 or FP32_MID_FP16. */
    int64_t workspace_limit;
} mfft_conv2d_desc;

typedef struct mfft_conv2d_s *mfft_conv2d;

typedef struct mfft_conv2d_info {
    int pad_h, pad_w, out_h, out_w;
    int64_t device_bytes;       /* resident spectra + buffers + plan workspaces */
    int64_t spectrum_bytes;     /* the resident filter spectra alone */
} mfft_conv2d_info;

int mfft_conv2d_desc_default(mfft_conv2d_desc *d);
int mfft_conv2d_create(mfft_conv2d *op, const mfft_conv2d_desc *d);
/* filters: device array [F][C][fh][fw] of MFFT_REAL_*; transforms and keeps the spectra */
int mfft_conv2d_set_filters(mfft_conv2d op, const void *filters, int real_fmt, void *stream);
/* in: device [B][C][h][w], out: device [B][F][out_h][out_w] (MFFT_REAL_* each) */
int mfft_conv2d_forward(mfft_conv2d op, const void *in, int in_fmt, void *out, int out_fmt, void *stream);
/* as forward, with only resident filter `filter_index` (out: [B][1][out_h][out_w]) */
int mfft_conv2d_forward_select(mfft_conv2d op, const void *in, int in_fmt, void *out, int out_fmt, int filter_index,
                               void *stream);
/* Two single-image convolutions in one complex transform: image in0 with filter filter0
 * and image in1 with filter filter1 (operators with channels = 1, any batch). The images
 * become the real and imaginary parts of one complex image; the pointwise step separates
 * their spectra (A = (X(k) + conj X(-k))/2, B = (X(k) - conj X(-k))/2i) and forms
 * A H0 + i B H1, so the inverse holds both results. Half the transform work of two
 * mfft_conv2d_forward_select calls. No counterpart in the reference (its batches pair
 * nothing). */
int mfft_conv2d_forward_two(mfft_conv2d op, const void *in0, const void *in1, int in_fmt, void *out0, void *out1,
                            int out_fmt, int filter0, int filter1, void *stream);
/* Gradients of the 'valid' correlation operator (MFFT_CONV_VALID_CORR; the reference's
 * fft_conv2d_half::conv2d_backward, tensorcore_fft_conv2d.cpp:364-442). delta = dL/dout,
 * [B][F][out_h][out_w]; in = the forward input [B][C][h][w] (needed for grad_w).
 *   grad_in [B][C][h][w]  = sum_f delta[b][f] 'full'-convolved with filter[f][c]
 *   grad_w  [F][C][fh][fw] = sum_b in[b][c] 'valid'-correlated with delta[b][f]
 * Either output may be NULL. All real buffers in real_fmt. The first call builds the
 * backward plans and a filter-gradient buffer (allocation); later calls allocate nothing. */
int mfft_conv2d_backward(mfft_conv2d op, const void *in, const void *delta, int real_fmt, void *grad_in, void *grad_w,
                         void *stream);
int mfft_conv2d_info_get(mfft_conv2d op, mfft_conv2d_info *info);
int mfft_conv2d_destroy(mfft_conv2d op);
/* smallest transform length >= n that runs on the fast path: 2^a * {1, 3, 5, 9, 15, 17} */
int mfft_good_size(int n);

/* ------------------------------------------------------------------------------------
 * Schedule autotuner. The reference fixes one schedule per size (tcfftCreate,
 * tcfft_half.cu:1250-1330; tcfftCreate_2d, tcfft_half_2d.cu:1167-1206); mfft_autotune times
 * the planner's candidates on the device -- four-step split point (split_hint) and batch
 * chunk (workspace_limit); never the precision or mid_fmt -- and caches the fastest per
 * (device arch, nx, ny, batch, precision, in/out/mid formats, direction, input order).
 * Plans created with split_hint = MFFT_SPLIT_TUNED use the cached schedule. ny = 0 for 1D.
 * The cache is process-wide; save/load keep it across processes (text, one entry a line).
 * ------------------------------------------------------------------------------------ */
#define MFFT_SPLIT_TUNED (-1)

typedef struct mfft_tune_entry {
    int split_hint;
    int64_t workspace_limit;
    double ms;           /* tuned schedule, ms per exec */
    double default_ms;   /* the planner's default schedule, ms per exec (-1 if it did not run) */
    int candidates;      /* schedules timed */
} mfft_tune_entry;

int mfft_autotune(int nx, int ny, int64_t batch, const mfft_desc *d, int reps, mfft_tune_entry *out);
int mfft_tune_lookup(const char *arch, int nx, int ny, int64_t batch, const mfft_desc *d, mfft_tune_entry *out);
int mfft_tune_cache_save(const char *path);
int mfft_tune_cache_load(const char *path);
int mfft_tune_cache_clear(void);
const char *mfft_tune_last_error(void);

/* ------------------------------------------------------------------------------------
 * Continuous-batching request server (no counterpart in the reference). Mixed 1D FFT,
 * 2D FFT and conv2d requests on device buffers are submitted without blocking; a
 * scheduler admits them while the in-flight device footprint fits `hbm_budget`, runs
 * admitted work concurrently on `streams` HIP streams, coalesces small 1D transforms
 * (n <= 4096) into one work-queue launch, and keeps plans, twiddles and filter spectra
 * resident; conv requests admitted together on one operator run in pairs, two images
 * in one complex transform (mfft_conv2d_forward_two). plain = 1 gives the baseline: fp32, one request at a time, each completed
 * before the next starts.
 * FFT requests: in/out are interleaved complex in the server's format (fp16 for split /
 * fp16 precision, fp32 for plain); out = out_scale * DFT(in) (mfft_completion.out_scale).
 * conv2d requests: in/out are fp32 real [nx][ny] images ('same' convolution with the
 * registered filter `filter_id`). Filters are registered before the first request; those
 * whose sizes have the same parity share one operator per image shape (each filter centred
 * in the class's largest frame), so all of their spectra stay resident together.
 * ------------------------------------------------------------------------------------ */
enum { MFFT_REQ_FFT1D = 1, MFFT_REQ_FFT2D = 2, MFFT_REQ_CONV2D = 3 };

typedef struct mfft_server_desc {
    int64_t hbm_budget;    /* bytes: workspace slabs + conv operators + in-flight request I/O */
    int streams;           /* concurrent HIP streams (1..16) */
    int precision;         /* MFFT_PREC_* (ignored in plain mode: fp32) */
    int max_batch_1d, max_batch_2d;
    int coalesce_max;      /* small 1D requests per work-queue launch */
    int64_t slab_bytes;    /* workspace per stream (default 512 MiB) */
    int plain;
    int mid_fmt;           /* intermediates of the FFT plans (mfft_desc.mid_fmt; -1 = default);
                              MFFT_FMT_C16 also gives the conv operators fp16 spectra */
} mfft_server_desc;

typedef struct mfft_request {
    int kind;              /* MFFT_REQ_* */
    int direction;         /* MFFT_FORWARD (default) / MFFT_INVERSE; FFT requests */
    int n;                 /* 1D length */
    int nx, ny;            /* 2D shape / conv image (nx rows of ny) */
    int64_t batch;         /* FFT transforms in the request */
    const void *in;
    void *out;
    int filter_id;         /* conv2d */
    int64_t tag;
} mfft_request;

typedef struct mfft_completion {
    int64_t tag;
    int status;            /* MFFT_OK or an error (MFFT_ERR_NOMEM: larger than the budget) */
    double submit_us, start_us, end_us;   /* host steady clock */
    double out_scale;
} mfft_completion;

typedef struct mfft_server_stats {
    int64_t submitted, completed, launches, coalesced;
    int64_t static_bytes, inflight_bytes;
    int64_t peak_bytes;    /* largest static + in-flight footprint at any admission */
    int64_t paired;        /* conv requests that shared a transform with another */
    int64_t busy_us;       /* scheduler thread time spent retiring, admitting and launching
                              (waits and sleeps excluded): busy_us / elapsed is its load */
} mfft_server_stats;

typedef struct mfft_server_s *mfft_server;
int mfft_server_desc_default(mfft_server_desc *d);
int mfft_server_create(mfft_server *server, const mfft_server_desc *d);
int mfft_server_add_filter(mfft_server server, int filter_id, const float *weights, int fh, int fw);
int mfft_server_submit(mfft_server server, const mfft_request *r);
int mfft_server_poll(mfft_server server, mfft_completion *out, int max, int *n);
int mfft_server_wait_idle(mfft_server server, double timeout_s);
int mfft_server_get_stats(mfft_server server, mfft_server_stats *stats);
int mfft_server_destroy(mfft_server server);
const char *mfft_server_last_error(void);

/* ------------------------------------------------------------------------------------
 * Launch trace of the calling thread (per-kernel timing of composite calls such as a conv
 * forward; no counterpart in the reference). Between mfft_trace_begin and mfft_trace_end
 * every kernel the library launches from this thread first records events[k] (hipEvent_t)
 * on its stream and fills records[k] with the kernel family and the algorithmic bytes it
 * reads and writes (each input and output element once); mfft_trace_end records
 * events[n] on `stream` after the last launch and returns n. Launch k then ran between
 * events[k] and events[k + 1] when every launch went to one stream. Launches beyond
 * `capacity` (nevents = capacity + 1 events) are not recorded and make mfft_trace_end
 * return MFFT_ERR_INVALID.
 * ------------------------------------------------------------------------------------ */
typedef struct mfft_launch_record {
    char kernel[64];        /* kernel family, line length and formats, e.g. "fast_cols R256 c16>c16" */
    int64_t bytes_read, bytes_written;
} mfft_launch_record;

int mfft_trace_begin(void **events, mfft_launch_record *records, int capacity);
int mfft_trace_end(void *stream, int *nlaunches);

/* Message of the calling thread's last error ("" if none). */
const char *mfft_last_error(void);

/* Library version string and the device's properties as seen by the library. */
const char *mfft_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MFFT_H */
