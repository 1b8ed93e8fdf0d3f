/*
 * CPU oracle for the adjoint angle-domain Kirchhoff migration of
 * kirchmig2d_cig (adj=1).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker or as the timed CPU baseline.
 *
 * Two builds of one restatement:
 *   *_f32 : the reference's own arithmetic (float storage, the double
 *           intermediates the reference uses in the Hermite interpolation and
 *           in the Kirchhoff pick), loop for loop, so that its CIG is bitwise
 *           identical to the reference binary built with g++ -O3 on x86-64;
 *   *_f64 : every value promoted to double and every operation done in
 *           double ("FP64 rerun of the reference's own summation").
 *
 * Parity pinning (which check fixes which function) is listed at the top of
 * kmig_oracle.c.
 */
#ifndef KMIG_ORACLE_H
#define KMIG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of one migration run, already parsed from key=value form.
   Field names follow the reference's variables in src/kirchmig2d_cig.cpp. */
typedef struct kmo_params {
    int adj;                  /* 1 migration (kmo_migrate), 0 modelling (kmo_model) */
    int cig;                  /* 0: stacked image, 1: CIG */
    int cig_is_angle;         /* mode=angle */
    int cmp;                  /* 1: data axis 2 is offset, 0: absolute receiver */
    int half_angle;
    float oazmin, oazmax, oaztaper;
    int isdiff;
    int aperture_trace;       /* -1: off */
    int aperture_trace_taper;
    int direct_mute, direct_mute_invert, direct_mute_qc_shot;
    float direct_mute_slope, direct_mute_intercept, direct_mute_taper;
    float tau, aperture, aperture_taper, aperture_rmin, aperture_rtaper, antialias;
    int angle_interp;
    float angle_taper_start, angle_taper_end;
    int angle_grad_radius, angle_smooth;
    int interp_type;          /* 'h', 'l' or 'p' */
    float g0, dg;             /* output CIG axis */
    int ng;
} kmo_params;

/* Axes of the data (t,h,s), the tables (z,x,pos) and their position axes. */
typedef struct kmo_geom {
    int nt, nh, ns;
    float t0, dt, h0, dh, s0, ds;
    int nz, nx;
    float z0, dz, x0, dx;
    int sny;
    float sy0, sdy;
    int rny;
    float ry0, rdy;
} kmo_geom;

/* Full migration over shots [shot_begin, shot_end).  data is (nt,nh,ns) for
   all shots; tables are (nz,nx,sny) and (nz,nx,rny); the four gradient
   volumes are either all NULL or all given.  out has ng*nz*nx elements and is
   zeroed here.  qc_before/qc_after (nt*nh each) may be NULL.
   Returns 0, or a negative code on invalid input. */
int kmo_migrate_f32(const kmo_params *p, const kmo_geom *g, const float *data,
                    const float *stab, const float *sder,
                    const float *rtab, const float *rder,
                    const float *sgx, const float *sgz,
                    const float *rgx, const float *rgz,
                    int shot_begin, int shot_end,
                    float *out, float *qc_before, float *qc_after);
int kmo_migrate_f64(const kmo_params *p, const kmo_geom *g, const float *data,
                    const float *stab, const float *sder,
                    const float *rtab, const float *rder,
                    const float *sgx, const float *sgz,
                    const float *rgx, const float *rgz,
                    int shot_begin, int shot_end,
                    double *out, float *qc_before, float *qc_after);

/* adj=0 modelling: image [ng][nx][nz] (float, as read from the reference's
   migration= file) to data [ns][nh][nt] (adj=0 outputs every trace; traces
   outside the trace aperture or the offset axis are zeros). */
void kmo_pick0_f64(int nt, double t0, double dt, double ti, double deltat, double sample,
                   double *trace);
int kmo_model_f32(const kmo_params *p, const kmo_geom *g, const float *image,
                  const float *stab, const float *sder, const float *rtab, const float *rder,
                  const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                  int shot_begin, int shot_end, float *data);
int kmo_model_f64(const kmo_params *p, const kmo_geom *g, const float *image,
                  const float *stab, const float *sder, const float *rtab, const float *rder,
                  const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                  int shot_begin, int shot_end, double *data);

/* Building blocks, exported so that tests can pin each one separately. */
void kmo_doubint_f32(int nt, float *trace);
void kmo_doubint_f64(int nt, double *trace);
/* adj=1 pick: returns the value added to a zero sample. */
float kmo_pick_f32(int nt, float t0, float dt, float ti, float deltat, const float *trace);
double kmo_pick_f64(int nt, double t0, double dt, double ti, double deltat, const double *trace);
/* Interpolate one table side at coordinate `pos`: writes T and dT/dpos.
   Also returns the clamped lower node index through *inode. */
void kmo_interp_side_f32(int type, int nz, int nx, float dx_grid, int npos, float o,
                         float d, float pos, const float *tab, const float *der,
                         float *T, float *Tx, int *inode);
void kmo_grad_plane_f32(const float *tab, int nz, int nx, float dz, float dx,
                        int radius, float *gx, float *gz);
void kmo_smooth_f32(float *f, int nz, int nx, int r);
float kmo_angle_f32(float sx, float sz, float rx, float rz, int half_angle);
int kmo_linear_bins_f32(float value, float g0, float dg, int ng,
                        int *ig0, int *ig1, float *w0, float *w1);
void kmo_mute_f32(float *trace, int nt, float t0, float dt, float sx, float rx,
                  float slope, float intercept, float taper, int inverted);
float kmo_leg_weight_f32(float leg_deg, float aper, float band);
float kmo_floor_weight_f32(float d, float rmin, float rt);
float kmo_depth_mute_f32(float angle, int iz, int nz, float amin, float amax, float band);
float kmo_angle_taper_f32(float angle, float start, float end);

#ifdef __cplusplus
}
#endif
#endif
