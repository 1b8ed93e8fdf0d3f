// Internal declarations shared by the host orchestration (kmig_api.cpp) and
// the HIP kernels (kmig_kernels.hip).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>

#include "kmig.h"

namespace kmig {

void set_err(char *err, int errlen, const char *fmt, ...);
int collect_kv(std::map<std::string, std::string> &kv, int argc, const char *const *argv,
               char *err, int errlen);
int parse_kv(kmig_params *p, const std::map<std::string, std::string> &kv, int nh, float h0,
             float dh, char *err, int errlen);

// Per-trace record written by the host from the acquisition geometry
// (src/kirchmig2d_cig.cpp:1055-1105): trace-aperture weight (negative = the
// trace is skipped), receiver tuple slot, offset-CIG bin, mute geometry.
struct TraceMeta {
    float wap;        // trace_ap_weight, or -1 when the trace does not migrate
    int slot;         // receiver tuple slot of this trace within its batch
    int ig;           // output bin of an offset CIG / stack (0)
    float src_x;      // source x (mute)
    float rcv_x;      // receiver x (mute)
    int pad[3];
};

// How one tuple slot is read from its side's table volume
// (src/kirchmig2d_cig.cpp:987-1028 / :1148-1180, tinterpsr.cpp:37-188).
struct SlotSpec {
    int node;         // lower table node (or the clamped node)
    int interp;       // 0: copy node, 1: interpolate node..node+1
    int k1, k2;       // 'p' shifts
    float coord;      // in-interval coordinate (float, as the reference)
    float t;          // coord / d
    float surf_x;     // surface position (aperture cone vertex)
    int gnode;        // gradient-volume lower node (linear, clamped)
    int ginterp;      // 0: copy, 1: interpolate
    float gw;         // gradient-volume weight
    int pad[2];
    double h[8];      // h00, h10*d, h01, h11*d, d00, d10*d, d01, d11*d
};

// Structure-of-arrays tuple set: one nz*nx plane per slot and field.
struct Tuples {
    float *T = nullptr;   // traveltime at the slot position
    float *A = nullptr;   // |dT/dpos * spacing| (anti-alias term)
    float *GX = nullptr;  // smoothed dT/dx
    float *GZ = nullptr;  // smoothed dT/dz
    float *W = nullptr;   // aperture leg weight (0 also where |grad| <= 1e-12)
    int cap = 0;          // allocated slots
};

struct TuplePrepArgs {
    const float *tab, *der, *gxv, *gzv;  // side volumes (npos*nzx)
    const SlotSpec *spec;
    int nslots;
    int nz, nx;
    size_t nzx;
    float x0, dx, dz, dd;                // image grid, dd = x spacing for 'p'
    float dpos;                          // table position spacing
    float dside;                         // ds (source) or dh (receiver)
    int type;                            // 'h', 'l', 'p'
    int use_grad_files;
    int radius, smooth;
    int angle_mode;
    float aper, aper_taper, rmin, rtaper;
    float aal;                           // antialias (folded into packed FP32 tuples)
    // fused pass only: every slot copies its table node (no interpolation),
    // so T and |dT/ds * dside| are read from tab / der at the slot's node
    // instead of the interpolation's planes (which are then not written)
    int direct;
    // with direct and a surrogate (per-slot fallback): the slots whose
    // largest residual nn_maxres[slot] is within nn_lim (float bits) took
    // the network, whose T and A are in the planes; the others read the node
    const unsigned *nn_maxres;
    unsigned nn_lim;
};

struct TracePrepArgs {
    const float *raw;
    float *prep;
    const TraceMeta *meta;
    float *qcb, *qca;
    int nt, nh;
    int shot_begin, shot_end;
    float t0, dt;
    int mute, invert, qc_shot, diff;
    float slope, intercept, taper;
};

struct MigArgs {
    int nz, nx;
    size_t nzx;
    int nt;
    float t0, dt, inv_dt;
    int nh, shot0, nshots;
    const float *sT, *sA, *sGX, *sGZ, *sW;
    const float *rT, *rA, *rGX, *rGZ, *rW;
    const TraceMeta *meta;
    const float *traces;
    float *cig;
    float *spray;        // adj=0: traces the image is sprayed into (cig is read only)
    float tau, aal;
    int mode;            // 0 angle/linear bins, 1 angle/nearest bin, 2 offset CIG or stack
    int half_angle;
    float g0, dg;
    int ng;
    float oazmin, oazmax, oaztaper;
    float taper_start, taper_end;
};

// Packed 16-byte tuple of the resident migrate kernel, one per (slot, point)
// (FP32 fields; precision=fp16 stores IEEE half fields in the same 16 bytes,
// described at tuple_h in kmig_prep.hip):
//   frac  fractional part of T/dt (samples, in [0,1])
//   a     1 + antialias * |dT/dpos * spacing| / dt: one plus the anti-alias
//         half-width in samples (the outer taps of an update lie 1 + antialias
//         * max(a_s, a_r) samples from its time, and that is fmaf(a, aal, 1) of
//         the larger side, so the larger folded field is it bit for bit)
//   th    direction of the smoothed traveltime gradient, atan2(gx, gz), in
//         output-angle degrees (x 180/pi, or x 90/pi for half opening angles)
//   iw    bits 31..16: integer part of T/dt (int16), bits 15..0: the
//         aperture weight W * 32768 (W in [0,1], 0 where |grad| <= 1e-12)
struct PackedTuple {
    float frac, a, th;
    uint32_t iw;
};

// 12-byte tuple of the BF16 and UNORM16 precision modes:
//   iw  as PackedTuple::iw (int16 integer sample | W * 2^15)
//   fa  bits 15..0: fractional sample part f in [0,1)
//       bits 31..16: a / a_scale of the slot, in [0,1]
//   th  gradient direction in FP32: the opening-angle bins need ~20 bits of it
//       (16-bit directions, even UNORM16 over the slot's range, leave CIG errors
//       of 2e-3 .. 8e-2 relL2 against the FP64 oracle)
// UNORM16 stores f as round(f * 65536) (65535 max) and a/a_scale as
// round(x * 65535).
struct PackedTuple16 {
    uint32_t iw, fa, th;
};

// Per-slot scale of the 16-bit anti-alias field (its maximum over the slot).
struct SlotScale {
    float a_scale;
};

// Bounds of the packed directions th (output-angle degrees) of one slot over
// one cell of BND_CZ x BND_CX image points, counting only points of nonzero
// tuple weight: lo = floor(min th * BND_Q), hi = ceil(max th * BND_Q), exact
// integers (|th| <= 180); tmin = the smallest integer sample (iw >> 16) of
// those points.  A cell with no such point has lo = tmin = BND_EMPTY > hi.
struct ThBound {
    int lo, hi;
    int tmin, pad;
};
constexpr int BND_CZ = 16, BND_CX = 4;
constexpr float BND_Q = 65536.0f;
constexpr int BND_EMPTY = 1 << 28;


// Bounds of every (slot, cell) of FP32 (pk) or 16-bit (pk16) packed tuples,
// cell-major: out[cell * nslots + slot], cell = cx * ncz + cz.
int launch_th_bounds(const PackedTuple *pk, const PackedTuple16 *pk16, int nslots, int nz,
                     int nx, ThBound *out, void *stream);

// One trace of a receiver column (resident kernel, receivers that move with
// the shot): its receiver index in the shot and its trace-aperture weight,
// or (0, -1) where the shot has no trace in the column.
struct ColTrace {
    int ih;
    float wap;
};

struct MigResArgs {
    int nz, nx;
    size_t nzx;
    int nt, nh, shot0, nshots;
    const PackedTuple *stup;     // [nshots][nzx] (batch-local shots), FP32 mode
    const PackedTuple *rtup;     // [slots][nzx]
    const PackedTuple16 *stup16; // the same in the 16-bit modes
    const PackedTuple16 *rtup16;
    const SlotScale *sscale;     // [nshots] (batch-local), 16-bit modes
    const SlotScale *rscale;     // [slots]
    const int *rslot;            // [nh + RSLOT_PAD] receiver slot of each receiver index
                                 // (or column), -1: unused (and past the last)
    const float *wap;            // [ns][nh] trace-aperture weight, < 0: trace skipped
    const float *pairs;          // [ns][nh][nt][2] (tr[k], tr[k+1] - tr[k])
    const ThBound *sbnd;         // [ncz * ncx][sbnd_stride] direction bounds of the
                                 // shots (entry js of a row: batch-local shot js),
                                 // or null (no culling)
    const ThBound *rbnd;         // [ncz * ncx][nrslots]
    int scell_off;               // batch index of local shot 0
    // Survivor masks of the group culls, written at the start of every launch
    // (k_cull_mask) from sbnd and rbnd: bit j of word w of (cell, receiver
    // chunk k) is set where launch-local shot 64 w + j may add a nonzero
    // update at a point of the cell with a receiver of chunk k
    // (idx::mask_word; chunks of the launch preset's R receivers, ceil(nshots
    // / 64) words).  Null: no culling.
    unsigned long long *mask;
    int ncz, ncx;                // cells of BND_CZ x BND_CX points
    int sbnd_stride, nrslots;    // shots per sbnd row, receiver slots
    float *spray;                // adj=0: [ns][nh][nt] traces the image is sprayed into
    float *cig;
    // Shot groups: the launch runs nsplit blocks per tile, group sg taking
    // batch-local shots [sg nshots / nsplit, (sg + 1) nshots / nsplit) and
    // writing its partial CIG to part[sg] ([nsplit][ng][nzx], group 0 starting
    // from the CIG); the partials are then summed into cig in group order.
    // part_splits = groups part holds (0: no buffer, one group).
    float *part;
    int part_splits;
    // counting launches only (launch_count_resident): [0] updates evaluated
    // (valid points x real receivers of every (shot, receiver chunk) group
    // that survives the culls), [1] updates accepted (taps inside the trace
    // and nonzero weight); the counting kernel writes no CIG
    unsigned long long *counts;
    // Receivers that move with the shot (cmp=1; ctab non-null): the kernel
    // sweeps receiver columns instead of receiver indices.  Column c keeps
    // receiver tuple slot rslot[c] for every shot; ctab[(scell_off + js) *
    // ncols + c] is the trace of batch shot scell_off + js in that column;
    // cfirst[c] .. clast[c] are the first and last batch shots with a trace
    // in it (clast < cfirst: none).  ctab is padded past its last row.
    const ColTrace *ctab;
    const int *cfirst, *clast;
    int ncols;
    // offset gathers and the stack (cig=0) on the resident kernel: the output
    // bin of each receiver index (offset bin, or 0 for the stack); null for
    // angle gathers
    const int *rbin;
    int c_int;                   // integer part of (tau - t0)/dt
    int t_cull;                  // nt - 2 - c_int: a group whose smallest source plus
                                 // receiver integer samples reach it has every
                                 // update's last tap past the trace
    float c_frac;                // fractional part of (tau - t0)/dt
    float aal;
    float wrap;                  // 360 (full opening angle) or 180 (half)
    // Bin coordinate u = fmaf(angle, inv_dg, g0_inv_dg), g0_inv_dg = -g0/dg.
    // u is non-decreasing in the angle (dg > 0), so u <= ng - 1 + 1e-5 holds
    // exactly for angle <= ang_hi (ang_hi_up: the next float above), and
    // u >= -1e-5 for angle >= ang_lo (tested only when lo_check; otherwise
    // every angle in [0, wrap / 2] passes).
    float inv_dg, g0_inv_dg;
    float ang_hi, ang_hi_up, ang_lo;
    int lo_check;
    int ng;
    float oazmin, oazmax, oaztaper, inv_oaztaper;
    float taper_start, taper_end, inv_taper;
};

// wap entries past the last trace: at least the largest receiver chunk (R)
constexpr int RES_WAP_PAD = 64;

// rslot entries (-1) past the last receiver: at least the largest chunk (R)
constexpr int RSLOT_PAD = 16;

// Most shot groups of one resident launch.
constexpr int KMIG_MAX_SPLIT = 16;

constexpr int KMIG_MAX_PROBLEMS = 8;

// Depth-layer stitch of angle gathers (layercom_cig): layer il covers full
// depth samples [start[il], end[il]] with CIG layer[il] ([ng][nx][n1_il]).
struct StitchArgs {
    const float *layer[KMIG_MAX_PROBLEMS];
    int start[KMIG_MAX_PROBLEMS], end[KMIG_MAX_PROBLEMS];
    float scale[KMIG_MAX_PROBLEMS];
    const float *lat;   // [nx] lateral edge scale (layers after the first)
    int nlayers, nz, nx, g0, ng_out;
    int stitch_blend;   // -1: blend over the whole overlap
};

struct MigMultiArgs {
    int nprob;
    int tile0[KMIG_MAX_PROBLEMS + 1];  // first block of each problem; tile0[nprob] = grid
    MigResArgs prob[KMIG_MAX_PROBLEMS];
};

// Elements of the raw, prepared and pair buffers for ntraces traces of nt
// samples: TRACE_PAD traces and 64 samples more, never written and never
// used (the trace preparation loads a block's 64 rows of a chunk of 32
// samples unclamped: from a block's first trace at most 64 rows plus 31
// samples past the last valid trace's end).
constexpr int TRACE_PAD = 64;
inline size_t trace_alloc(size_t ntraces, int nt)
{
    return (ntraces + TRACE_PAD) * (size_t)nt + 64;
}

// Launchers (kmig_prep.hip, kmig_migrate.hip).  Return 0 or a hipError_t value.
int launch_trace_prep(const TracePrepArgs &a, void *stream);
// coalesced trace preparation; pairs (nullable) gets (tr[k], tr[k+1] - tr[k]); prep
// keeps the prepared samples only where pairs is null or diff=1 (otherwise it is
// left holding the causal sums); tmp is a trace-sized scratch volume used only
// when diff=1
int launch_trace_prep2(const TracePrepArgs &a, float *pairs, float *tmp, void *stream);
int launch_tuple_prep(const TuplePrepArgs &a, Tuples &out, float *tmp1, float *tmp2,
                      void *stream);
// the same in two parts: table interpolation (T, A, gradient files), then
// gradients, smoothing and weights
int launch_tuple_interp(const TuplePrepArgs &a, Tuples &out, void *stream);
int launch_tuple_finish(const TuplePrepArgs &a, Tuples &out, float *tmp1, float *tmp2,
                        void *stream);
// gradients, smoothing, weights and packing into the 16-byte layout (FP32 or
// FP16 fields: precision) in one LDS-tiled pass (angle mode, radius and
// smooth <= 4); the packed tuples equal those of launch_tuple_finish +
// launch_pack bit for bit
// With bnd set the pass also writes the slots' direction and sample bounds
// (launch_th_bounds' output for the same tuples).
bool tuple_fused_ok(const TuplePrepArgs &a);
int launch_tuple_fused(const TuplePrepArgs &a, Tuples &o, PackedTuple *out, float dt,
                       float th_scale, int precision, ThBound *bnd, void *stream);
// the same for the 12-byte formats (BF16, UNORM16); the per-slot anti-alias scale is the
// maximum over all points of the slot (the unfused path takes it over the
// points with W > 0)
int launch_tuple_fused16(const TuplePrepArgs &a, Tuples &o, PackedTuple16 *out,
                         SlotScale *scale, float dt, float th_scale, int precision,
                         ThBound *bnd, void *stream);
int launch_migrate(const MigArgs &a, int tile, void *stream);
// adj=0 with the same kernel: a.cig is the image, a.spray the zeroed traces
int launch_model(const MigArgs &a, int tile, void *stream);
int migrate_lds_bytes(int ng, int tile);
int launch_fill(float *p, size_t n, float v, void *stream);
// *bad = 1 where a traveltime is non-finite or -0.0 or a derivative non-finite
int launch_table_check(const float *T, const float *D, size_t n, unsigned *bad, void *stream);
// th_scale: degrees per radian of the packed direction (0: no direction);
// precision: KMIG_PREC_FP32 or KMIG_PREC_FP16 fields
int launch_pack(const Tuples &in, PackedTuple *out, int nslots, size_t nzx, float dt,
                float th_scale, float aal, int precision, void *stream);
int launch_pack16(const Tuples &in, PackedTuple16 *out, SlotScale *scale, int nslots,
                  size_t nzx, float dt, float th_scale, int precision, void *stream);
int launch_pairs(const float *prep, float *pairs, long ntraces, int nt, void *stream);
int launch_stitch(const StitchArgs &a, float *out, unsigned *nonfinite, void *stream);
// the same with the lateral weights and the counter in the library's static
// device scratch (no allocation); waits for the stream and returns the count
// of non-finite samples in *nonfinite_host
constexpr int KMIG_MAX_DEVICES = 64;
int launch_stitch_noalloc(StitchArgs a, int scalex_aper, double scalex_scale, float *out,
                          unsigned *nonfinite_host, void *stream);
// tile: preset 9..21 (kmig_migrate.hip); precision: KMIG_PREC_* tuple format
int launch_migrate_resident(const MigResArgs &a, int tile, int precision, void *stream);
// adj=0: sprays the image held in a.cig into a.spray (FP32 tuples; presets
// 9, 11 and 15)
// Trace window (samples) one wave of the resident model kernel sums in LDS.
constexpr int MODEL_WIN = 256;
int launch_model_resident(const MigResArgs &a, int tile, void *stream);
int launch_migrate_multi(const MigMultiArgs &m, int tile, int precision, void *stream);
int migrate_tiles(int nz, int nx, int tile);
// the same kernels under a distinct symbol, for the autotuner's timing launches
int launch_migrate_resident_probe(const MigResArgs &a, int tile, int precision, void *stream);
// the FP32 resident kernel of preset `tile` with the same grid and culls,
// counting evaluated and accepted updates into a.counts instead of writing
// the CIG
int launch_count_resident(const MigResArgs &a, int tile, void *stream);

}  // namespace kmig
