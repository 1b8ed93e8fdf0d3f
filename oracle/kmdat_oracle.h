/* CPU oracle for kirchdat2d_freq (frequency-domain Kirchhoff datuming, the
   direct operator path).  TEST INFRASTRUCTURE ONLY. */
#ifndef KMDAT_ORACLE_H
#define KMDAT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int cmp;          /* 1: h axis relative to the source, 0: absolute */
    int aper;         /* aperture= (traces) */
    int tap;          /* taper= (summation samples) */
    int data_tap;     /* data_taper= (traces at the acquisition edges) */
    float antialias;  /* antialias= */
    float length;     /* length= (filter length, s) */
    float sdatum, rdatum;
    int use_bf;       /* use_bf=: 1 applies each frequency's operator through the
                         segmented butterfly (f32 only) */
    int bf_p, bf_leaf, bf_levels;   /* bf_p=, bf_n_leaf=, bf_panel_levels= */
    float bf_amp_eps, bf_phase_tol; /* bf_amp_eps=, bf_phase_tol= */
} kmdo_cfg;

/* Full path: data (ns, nh, nt) -> out (ns, nh, nt); interm (the
   receiver-side result, not reversed) when not NULL.  Frequencies outside
   [iw_begin, iw_end) are left zero (iw_end <= 0: all).  use_bf selects the
   butterfly operators (kmdo_datum_f32 only; the f64 build returns -1 for it).
   Returns 0, -1 bad arguments, -2 "Receiver table too small.", -3 "Source
   table too small.". */
int kmdo_datum_f32(const kmdo_cfg *cfg, const float *data, int nt, int nh, int ns, float dt,
                   float h0, float dh, float s0, float ds, const float *stable, int nsg,
                   float sg0, float dsg, const float *rtable, int nrg, float rg0, float drg,
                   float *out, float *interm, int iw_begin, int iw_end);
int kmdo_datum_f64(const kmdo_cfg *cfg, const float *data, int nt, int nh, int ns, float dt,
                   float h0, float dh, float s0, float ds, const float *stable, int nsg,
                   float sg0, float dsg, const float *rtable, int nrg, float rg0, float drg,
                   float *out, float *interm, int iw_begin, int iw_end);

/* Pieces, for unit tests. */
int kmdo_nfft(int nt, int nsam, float dt, int nrg, int nsg, const float *rtable,
              const float *stable);
void kmdo_filter_f32(float tau, float dt, int nsam, float *F);
void kmdo_H_f32(float tau, float dt, int nsam, int nfft, int iw, float *hr, float *hi);
void kmdo_H_f64(double tau, double dt, int nsam, int nfft, int iw, double *hr, double *hi);
float kmdo_taper_weight(int left, int center, int right, int tap, int stride);
float kmdo_edge_taper(int i, int n, int tap);
float kmdo_antialias_weight(float antialias, int nfft, float dt, int iw, float dtau);
void kmdo_fft_f64(int n, double *re, double *im, int inverse);
/* greensol_auto: tgreen[i1][i2] = ttable[(i1 * nx + ix2) * nz + idatum] with
   ix2 the nearest x sample of position i2.  Returns 0, -1 datum out of range,
   -2 "Receiver table too small.". */
int kmdo_greensol(const float *ttable, int nz, int nx, int ngx, float oz, float dz, float ox,
                  float dx, float ogx, float dgx, float datum, float *tgreen);

#ifdef __cplusplus
}
#endif
#endif
