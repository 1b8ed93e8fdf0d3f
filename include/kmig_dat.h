/* libkmdat: frequency-domain Kirchhoff datuming on MI355X (gfx950).
 *
 * Replaces the computing part of the reference's kirchdat2d_freq
 * (src/kirchdat2d_frequency_cs.cpp:1177-1413): the r2c transform of the
 * time-reversed, edge-tapered traces (read_traces_to_freq, :281-313), the
 * receiver-side operator (:493-546 direct, :814-947 butterfly), the
 * source-side operator (:548-672 direct, :949-1152 butterfly) and the c2r
 * transform back to reversed traces (write_traces_from_freq, :470-491), plus
 * greensol_auto (src/greensol_auto.cpp:23-101), which cuts the datum table
 * out of a traveltime cube.
 *
 * use_bf=0 applies the operators exactly (the reference's direct sums).
 * use_bf=1 (the default) reproduces the reference's segmented butterfly
 * (SERECKIRCH/src/bf1d.cpp, :814-1121): per frequency the panels the
 * reference would factor -- of each side's matrix for cmp=0, of each shot's
 * and each receiver gather's for cmp=1 -- get the factor's own values (the
 * same Chebyshev nodes, Lagrange values, bilinear samples and switch
 * kernel), every other entry its exact value.
 *
 * Layouts are the reference's RSF ones: data and outputs (ns, nh, nt) with
 * time fastest, datum tables (n, n) with the output position fastest
 * (table[summation][output], as greensol_auto writes them).  All calls
 * return KMDAT_OK (0) or a negative code; kmdat_last_error gives the
 * message, which is the reference's wherever the reference has one.
 */
#ifndef KMIG_DAT_H
#define KMIG_DAT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMDAT_OK 0
#define KMDAT_E_PARAM -1
#define KMDAT_E_DIM -2
#define KMDAT_E_DEVICE -4
#define KMDAT_E_STATE -5

/* The keys of kirchdat2d_freq (:1216-1270) with their defaults. */
typedef struct {
    int cmp;               /* 1: h axis relative to the source (x_r = s + h); 0: absolute */
    int aperture;          /* summation aperture, traces (50) */
    int taper;             /* cosine taper at the aperture ends, summation samples (10) */
    int data_taper;        /* half-cosine taper of the acquisition edges, traces (0) */
    float antialias;       /* operator antialias factor, 0 disables it (1) */
    float length;          /* datuming filter length, s (0.025) */
    float sdatum, rdatum;  /* datum depths of the source and receiver tables */
    int use_bf;            /* (1) the butterfly panels, 0 exact sums */
    int bf_p, bf_n_leaf, bf_panel_levels;  /* (12, 16, 1) */
    float bf_amp_eps, bf_phase_tol;        /* (1e-20, 2) */
    int interm;            /* keep the receiver-side result (the interm= file) */
} kmdat_config;

typedef struct {
    int nfft, nw, nsam;
    long long keys_receiver, keys_source;    /* table keys (rows of the operator tables) */
    long long pairs_receiver, pairs_source;  /* summation pairs */
    double cmacs;           /* complex multiply-adds of both operators (pairs x nw) */
    float ms_forward;       /* pad + r2c */
    float ms_tables;        /* both operator tables */
    float ms_receiver, ms_source;  /* the two operators */
    float ms_inverse;       /* c2r + reversal (and the interm cube) */
    float ms_total;
    int banded;             /* 1: cmp=0 banded kernel, 0: pair lists */
    int butterfly;          /* 1: use_bf=1 ran the butterfly panels */
    long long bf_panels_receiver, bf_panels_source;  /* panels that can be butterflies */
    long long bf_items_receiver, bf_items_source;    /* (panel, 64-frequency chunk) items */
    float ms_bf_receiver, ms_bf_source;  /* the butterfly steps, inside ms_receiver / ms_source */
} kmdat_stats;

typedef struct kmdat_ctx kmdat_ctx;

void kmdat_config_default(kmdat_config *cfg);
/* The reference's checks of the keys (:1216-1261). */
int kmdat_config_check(const kmdat_config *cfg, char *err, size_t errlen);

int kmdat_create(kmdat_ctx **ctx, const kmdat_config *cfg);
void kmdat_destroy(kmdat_ctx *ctx);
const char *kmdat_last_error(const kmdat_ctx *ctx);

/* side 0: source table (sgreen_file), 1: receiver table (rgreen_file);
   table[n][n], positions o + k d (:1280-1298). */
int kmdat_set_green(kmdat_ctx *ctx, int side, const float *table, int n, float o, float d);
/* The (ns, nh, nt) cube (the input_file); copied to the device. */
int kmdat_set_data(kmdat_ctx *ctx, const float *data, int nt, float dt, int nh, float h0,
                   float dh, int ns, float s0, float ds);
/* Enqueue the whole datuming on the context's stream (asynchronous). */
int kmdat_run(kmdat_ctx *ctx);
int kmdat_sync(kmdat_ctx *ctx);
int kmdat_get_output(kmdat_ctx *ctx, float *out);        /* ns*nh*nt, reversed back */
int kmdat_get_interm(kmdat_ctx *ctx, float *interm);     /* needs cfg.interm */
int kmdat_get_stats(const kmdat_ctx *ctx, kmdat_stats *st);
/* One call: set_data + run + sync + get_output (+ interm when not NULL). */
int kmdat_datum(kmdat_ctx *ctx, const float *data, int nt, float dt, int nh, float h0, float dh,
                int ns, float s0, float ds, float *out, float *interm);

/* greensol_auto: tgreen[i1][i2] = T(position i1; x of position i2, datum
   depth) from the (nz, nx, ngx) cube ttable; returns KMDAT_OK or
   KMDAT_E_PARAM ("Datum out of range.") / KMDAT_E_DIM ("Receiver table too
   small."). Host only. */
int kmdat_greensol(const float *ttable, int nz, int nx, int ngx, float oz, float dz, float ox,
                   float dx, float ogx, float dgx, float datum, float *tgreen);

#ifdef __cplusplus
}
#endif
#endif
