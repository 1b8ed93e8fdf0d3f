/*
 * kmig.h — C ABI of the MI355X Kirchhoff angle-gather imaging library
 * (libkmig.so), a drop-in for the adjoint path of the reference's
 * kirchmig2d_cig program.
 *
 * What each entry point replaces in the reference:
 *   kmig_params_default / kmig_params_parse
 *        the key=value parsing, defaults and validation of main(),
 *        src/kirchmig2d_cig.cpp:548-875, with parse_output_axis :323-368
 *        and the se_get_par_* casts of SEFILESYSTEM/src/se_par_sep.cpp:432-446;
 *   kmig_create
 *        the one-time set-up of main() (:924-962: buffers, tinterp_init,
 *        kirmig_init; SERECKIRCH/src/tinterpsr.cpp:27-35, kirmig.cpp:26-32);
 *   kmig_set_tables
 *        the slice streamer load_table_slice / interp_volume_linear_stream
 *        (:274-322) and the table checks of :763-815 — the whole table is made
 *        resident in HBM instead of streamed two slices at a time;
 *   kmig_set_data
 *        the sequential trace reads se_fsio_read_float (:1069);
 *   kmig_run / kmig_sync
 *        the shot loop :979-1353 (tinterp_hermite/dinterp_hermite,
 *        compute_grad_plane, smooth_grad_component, apply_direct_wave_mute,
 *        doubint/diff2, the OpenMP image-point loop with kirmig_pick(adj=1));
 *   kmig_get_cig / kmig_get_qc
 *        the output writes se_fsio_write_float (:1071, :1080, :1359);
 *   kmig_migrate
 *        all of the above in one call;
 *   kmig_last_error
 *        the ERROR(...) message that the reference prints before exit(1)
 *        (SEBASIC/src/se_log.cpp:426-486); here every error is a status code.
 *
 * Conventions: 0 = OK, negative = error class (KMIG_E_*).  No entry point
 * exits the process.  One context per host thread; calls on one context are
 * serialised by the caller; all device work of a context runs on one HIP
 * stream owned by the context.  Array layouts follow RSF axis order with
 * axis 1 fastest: data (t,h,s) = data[(is*nh + ih)*nt + it], tables
 * (z,x,pos) = T[(ipos*nx + ix)*nz + iz], CIG (z,x,g) = cig[(ig*nx + ix)*nz + iz].
 */
#ifndef KMIG_H
#define KMIG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMIG_OK 0
#define KMIG_E_PARAM (-1)   /* invalid or missing parameter            */
#define KMIG_E_DIM (-2)     /* inconsistent array dimensions            */
#define KMIG_E_IO (-3)      /* file input/output (RSF helpers, CLI)     */
#define KMIG_E_DEVICE (-4)  /* HIP runtime error or no usable GPU       */
#define KMIG_E_STATE (-5)   /* call made before its inputs were set     */
#define KMIG_E_NOMEM (-6)   /* device or host allocation failed         */

/* Migrate kernel variants.  STREAM keeps the reference's per-point order
   (shot-major, receiver-minor) and per-update arithmetic; RESIDENT keeps a
   chunk of receiver tuples in registers and sweeps the shots under it
   (receivers that do not move with the shot, cmp=0), with the opening angle
   from per-side directions and split-sample traveltimes. */
#define KMIG_VARIANT_AUTO 0
#define KMIG_VARIANT_STREAM 1
#define KMIG_VARIANT_RESIDENT 2

/* Table precision of the resident traveltime tuples. */
#define KMIG_PREC_FP32 0
#define KMIG_PREC_FP16 1   /* 16-bit IEEE half fields, per-slot offset/scale   */
#define KMIG_PREC_BF16 2   /* 16-bit bfloat16 fields, per-slot offset/scale    */
#define KMIG_PREC_U16 3    /* 16-bit unsigned-normalised fields, same scaling  */

/* Every key of kirchmig2d_cig's adj=1 path, same names, units and defaults
   (src/kirchmig2d_cig.cpp:520-560, :568-598, :707-717, :816-862, :952). */
typedef struct kmig_params {
    int adj;                   /* 1 migration (kmig_run), 0 modelling (kmig_model) */
    int cig;                   /* 1: common-image gathers, 0: stacked image */
    int cig_is_angle;          /* mode=angle (1) or mode=offset (0)         */
    int cmp;                   /* 1: data axis 2 is offset, 0: receiver x   */
    int half_angle;            /* 0: full opening angle, 1: half opening    */
    float oazmin, oazmax, oaztaper; /* depth-varying angle mute (degrees)  */
    int diff;                  /* 1: centred derivative after doubint       */
    int aperture_trace;        /* -1: off; else max |iht - ist| (table idx) */
    int aperture_trace_taper;
    int direct_mute, direct_mute_invert, direct_mute_qc_shot;
    float direct_mute_slope, direct_mute_intercept, direct_mute_taper;
    float tau, aperture, aperture_taper, aperture_rmin, aperture_rtaper, antialias;
    int angle_interp;          /* 1: two-bin linear, 0: nearest bin         */
    float angle_taper_start, angle_taper_end;
    int angle_grad_radius, angle_smooth;
    int interp_type;           /* 'h' (Hermite), 'l' (linear), 'p' (partial) */
    float g0, dg;              /* output CIG axis (origin, step)            */
    int ng;                    /* output CIG axis length (1 if cig == 0)    */
    /* MI355X-side options (no counterpart keys in the reference) */
    int precision;             /* KMIG_PREC_*                               */
    int tile;                  /* migrate tile preset, 0 = automatic        */
    int variant;               /* KMIG_VARIANT_*                            */
} kmig_params;

typedef struct kmig_stats {
    double updates;            /* nominal ns*nh*nz*nx of the last run       */
    double ms_trace_prep;      /* kernel time, HIP events on the ctx stream */
    double ms_tuple_prep;
    double ms_migrate;
    double ms_total;           /* first to last event of the last run       */
    int migrate_launches;      /* migrate kernel launches in the last run   */
    int receiver_slots;        /* resident receiver tuples                  */
    int shot_batches;
    int variant;               /* KMIG_VARIANT_* actually used              */
    int tile;                  /* migrate tile preset used (autotuned if 0) */
    int graph;                 /* 1: the run replayed a captured step (only
                                  ms_total is its own; the phase times are
                                  the eager run's it was captured from)     */
    int tuple_direct;          /* tuple passes (side x shot batch) that read
                                  the table nodes directly: every slot sits
                                  on a node, so no interpolation pass ran   */
} kmig_stats;

typedef struct kmig_ctx kmig_ctx;

/* Fill *p with the reference's defaults (mode=offset, cig=0, cmp=1, ...). */
void kmig_params_default(kmig_params *p);

/* Apply key=value arguments with the reference's semantics: later keys
   override earlier ones, integers parse as (int)double and floats as
   (float)double; mode=angle turns cig on unless cig= is given; the output
   axis is derived from the data axis 2 (nh,h0,dh) and the cig_ / angle_ /
   offset_ keys.  File keys (seismic_data=, ...) are ignored here.
   Returns KMIG_OK or KMIG_E_PARAM with the reference's message in err. */
int kmig_params_parse(kmig_params *p, int argc, const char *const *argv, int nh, float h0,
                      float dh, char *err, int errlen);

/* Host-only validation of a parameter set against the data axes: the
   checks of src/kirchmig2d_cig.cpp:719-738 (direct_mute options and QC
   shot range), :359/:363 (axis) and :956-957 (interp type). */
int kmig_params_check(const kmig_params *p, int ns, char *err, int errlen);

/* Create a context on the current HIP device.  Validates p. */
int kmig_create(kmig_ctx **out, const kmig_params *p);
void kmig_destroy(kmig_ctx *ctx);
const char *kmig_last_error(const kmig_ctx *ctx);

/* Make one side's tables resident (side 0: source, 1: receiver).  T and
   dTds are (nz,nx,npos); gx/gz are the optional spatial-gradient volumes
   (either both NULL or both given on both sides before kmig_run).  Host
   pointers are borrowed for the duration of the call.  The receiver table
   must match the source table's nz and nx. */
int kmig_set_tables(kmig_ctx *ctx, int side, const float *T, const float *dTds,
                    const float *gx, const float *gz, int nz, int nx, int npos, float z0,
                    float dz, float x0, float dx, float o_pos, float d_pos);

/* Upload the seismic data (nt,nh,ns) to HBM with its axes. */
int kmig_set_data(kmig_ctx *ctx, const float *data, int nt, float t0, float dt, int nh,
                  float h0, float dh, int ns, float s0, float ds);

/* Migrate shots [shot_begin, shot_end) into the context's CIG, which is
   zeroed first.  Enqueues trace preparation, tuple preparation and the
   migration kernels on the context stream and returns without waiting. */
int kmig_run(kmig_ctx *ctx, int shot_begin, int shot_end);
int kmig_sync(kmig_ctx *ctx);

/* Captured steps (no counterpart in the reference, which has no device
   work): with enable = 1, a resident run (mode=angle, linear bins, one shot
   batch, no surrogate) is captured as one hipGraph right after it is
   enqueued eagerly, and later kmig_run calls with the same plan, shot range,
   preset and switches replay the graph instead of launching kernel by kernel:
   trace preparation (on a parallel branch), tuple preparation, the migrate
   launch and its partial sums.  Any change of tables, data, range or preset
   runs eagerly and captures again.  enable = 0 drops the graph.  Results are
   bitwise those of eager runs. */
int kmig_set_graph(kmig_ctx *ctx, int enable);

/* Copy results to host memory after kmig_sync: cig has ng*nx*nz floats;
   qc_before/qc_after have nt*nh floats (the direct_mute_qc_shot gather). */
int kmig_get_cig(kmig_ctx *ctx, float *cig);
int kmig_get_qc(kmig_ctx *ctx, float *qc_before, float *qc_after);
int kmig_get_stats(const kmig_ctx *ctx, kmig_stats *st);

/* How much of the nominal update grid of the last run survives the culls of
   the resident kernel (no counterpart in the reference, which evaluates every
   (shot, receiver, image point) update, src/kirchmig2d_cig.cpp:1203-1339).
   Re-runs the last run's preset with the same grid and culls, counting
   instead of accumulating (the CIG is untouched), and waits for it.
   nominal = ns*nh*nz*nx; evaluated = updates in (shot, receiver group)
   groups that passed the direction and time-window culls; accepted = updates
   whose taps lie inside the trace with a nonzero weight.  Needs a previous
   kmig_run on the resident kernel with FP32 tuples. */
int kmig_count_updates(kmig_ctx *ctx, double *nominal, double *evaluated, double *accepted);

/* One call: set_data + run(all shots) + sync + get_cig (+ QC when not NULL). */
int kmig_migrate(kmig_ctx *ctx, const float *data, int nt, float t0, float dt, int nh,
                 float h0, float dh, int ns, float s0, float ds, float *cig,
                 float *qc_before, float *qc_after);

/* adj=0 Kirchhoff modelling, replacing the adj=0 branch of
   src/kirchmig2d_cig.cpp:979-1351 (data axes from its nt/t0/dt/nh/h0/dh/ns/s0/
   ds keys, :663-701): the image (ng*nx*nz floats, the migration= file) is
   sprayed with the transposed kirmig_pick (SERECKIRCH/src/kirmig.cpp:59-67)
   and the migration's weights, bins and apertures into every trace, each
   trace is doubly integrated, and data receives ns*nh*nt floats (zeros
   outside the trace aperture).  It is the adjoint of kmig_run without mute
   and diff, for every output mode (angle, offset, stack) and cmp; the
   context needs adj=0 and precision=fp32.  Synchronous.  The reference's own adj=0 run ends in a segmentation fault
   after its axis message (the output header's ndim is never set, :875-920). */
int kmig_model(kmig_ctx *ctx, const float *image, int nt, float t0, float dt, int nh,
               float h0, float dh, int ns, float s0, float ds, float *data);

/* The HIP stream of the context (a hipStream_t), for callers that time or
   order their own work around kmig_run. */
void *kmig_stream(kmig_ctx *ctx);

/* ---- batches: several problems in one step ----
   A batch groups contexts whose tables and data are set (for example the
   depth layers of a recursive workflow, each with its own grid, shots,
   receivers and mute; no counterpart in the reference, which runs them as
   separate processes, run_salt_adcig.sh:171-181).  kmig_batch_run plans each
   context, then enqueues every context's trace and tuple preparation and ONE
   migrate launch covering all of them; the whole step is captured in a
   hipGraph on the first run and replayed afterwards.  Every context must use
   the resident kernel (mode=angle, angle_interp=1, cmp=0, one shot batch) and
   the same angle axis length and precision.  Call kmig_batch_sync before
   reading results with kmig_get_cig / kmig_get_qc of each context. */
typedef struct kmig_batch kmig_batch;
int kmig_batch_create(kmig_batch **out, kmig_ctx *const *ctx, int n);
int kmig_batch_run(kmig_batch *b);
int kmig_batch_sync(kmig_batch *b);
/* updates (sum of ns*nh*nz*nx) and the last step's time between HIP events
   recorded around the graph launch on the batch stream */
int kmig_batch_stats(const kmig_batch *b, double *updates, double *ms, int *graph_nodes);
const char *kmig_batch_last_error(const kmig_batch *b);
void kmig_batch_destroy(kmig_batch *b);
/* After kmig_batch_sync: the depth-layer stitch of the batch's angle gathers
   into one (z, x, angle) volume, replacing src/layercom_cig.cpp:412-557 (its
   checks and messages, :245-305 and :424-426) and the angle window of
   run_salt_adcig.sh:234-236.  The contexts are the layers, top to bottom;
   layer k covers full depth samples [start[k], end[k]] (its nz), each layer
   overlaps the one above and extends deeper.  scale (nullable: 1) multiplies
   layer k; layers after the first are also scaled by the lateral edge ramp
   scalex_aper / scalex_scale; the last stitch_blend overlap samples (-1: all)
   blend with complementary raised-cosine weights.  Angle samples
   [g_first, g_first + g_count) are written (g_count <= 0: all) to out
   (host, [g_count][nx][end[n-1] + 1]). */
int kmig_batch_stitch(kmig_batch *b, const int *start, const int *end, const float *scale,
                      int stitch_blend, int scalex_aper, double scalex_scale, int g_first,
                      int g_count, float *out);

/* Allocation-free form of kmig_batch_stitch, for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  No counterpart in the reference.  The
   same checks, weights and result as kmig_batch_stitch, bit for bit, except
   that out is DEVICE memory (caller-owned, on the layers' GPU, at least
   g_count x nx x (end[n-1] + 1) floats) and that the lateral edge weights and
   the non-finite counter live in a static device scratch array of the
   library instead of per-call hipMalloc buffers: the call makes no
   hipMalloc, hipFree or other allocating call.  Limit: nx <=
   KMIG_STITCH_NOALLOC_MAX_NX (KMIG_E_PARAM beyond).  The scratch array is
   shared by every call on one GPU and keeps the last call's contents, so
   calls that overlap on one GPU (two host threads, two batches) race.
   Synchronous: returns once out is written. */
#define KMIG_STITCH_NOALLOC_MAX_NX 65536
int kmig_batch_stitch_noalloc(kmig_batch *b, const int *start, const int *end,
                              const float *scale, int stitch_blend, int scalex_aper,
                              double scalex_scale, int g_first, int g_count, float *out);

/* ---- learned traveltime surrogate (include/kmig_nn.h) ----
   No counterpart in the reference, whose traveltimes always come from the
   tables (load_table_slice and tinterp_hermite, src/kirchmig2d_cig.cpp:274-284,
   :987-1029, :1148-1180).  With a network set on a side, every run evaluates
   it at each (slot position, grid point) right after the table interpolation
   and compares its traveltime with the table's: with the per-slot fallback
   (the default) a slot -- one shot position on the source side, one receiver
   position on the receiver side -- takes the network's traveltime and dT/ds
   at every point when its largest |T_nn - T_table| is within max_dt
   (seconds) and keeps its tables at every point otherwise; with the
   per-point fallback each point decides alone.  Gradients, smoothing,
   weights and the migration then proceed as without it.  nn = NULL removes
   it.  The network must outlive its use by the context. */
struct kmig_nn;
int kmig_set_surrogate(kmig_ctx *ctx, int side, struct kmig_nn *nn, float max_dt);
#define KMIG_NN_FALLBACK_SLOT 0
#define KMIG_NN_FALLBACK_POINT 1
int kmig_set_surrogate_fallback(kmig_ctx *ctx, int side, int mode);
/* After a run: grid points of that side's slots, points that took the
   network's values, and the largest |T_nn - T_table| seen.  With the
   per-slot fallback a slot whose residual already exceeds max_dt at a sample
   of its points (every 1021st) is not evaluated further, so for such a slot the
   residual seen is that of the sample: a lower bound, above max_dt. */
int kmig_get_surrogate_stats(kmig_ctx *ctx, int side, double *points, double *used,
                             float *max_residual);

/* Measurement aid, no counterpart in the reference: times `iters` rounds of
   64-lane 8-byte gathers from an L1-resident 16 KB buffer (every lane its own
   128-byte line for lanes_per_line = 1, or lanes_per_line consecutive lanes
   sharing one) over 2048 blocks of 256 threads.  Under a profiler its L1 and
   texture-data counters give the peak rate of the gathers the migrate
   kernels issue.  *ms = the timed launch, *wave_loads = its wave-level load
   instructions.  Allocates its own buffers. */
int kmig_probe_gather(int lanes_per_line, int iters, double *ms, double *wave_loads);

/* Library version string. */
const char *kmig_version(void);

#ifdef __cplusplus
}
#endif
#endif
