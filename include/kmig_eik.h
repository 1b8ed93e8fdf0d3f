/* Traveltime tables with source derivatives by fast marching (the eikods2d
 * table source), on the host.
 *
 * kmig_eikods2d replaces the computing part of the reference's eikods2d
 * program (src/eikods2d.cpp:150-765 with SERECKIRCH/src/eikods.cpp,
 * neighbors.cpp and pqueue.cpp): velocity in, per position the first-arrival
 * traveltime, its derivatives with respect to the relative (tdl1) and the
 * source (tds1) coordinate along axis l, and optionally the branch-aware
 * upwind spatial gradients (gradz, gradx).  The RSF reading and writing of
 * the program stay with the caller.  The outputs are the reference program's
 * bit for bit for 2-D models (n3 = 1; no plane-wave sources, no second-order
 * derivatives, fast marching rather than efmm).
 */
#ifndef KMIG_EIK_H
#define KMIG_EIK_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    KMIG_EIK_OK = 0,
    KMIG_EIK_W_FRONT = 1,  /* a front ran empty before every point was reached
                              (the reference warns "heap exausted" and goes on) */
    KMIG_EIK_E_PARAM = -1
};

typedef struct kmig_eik_config {
    int order;     /* accuracy order 1 or 2 (eikods2d order=, default 2) */
    int b1, b2;    /* constant-velocity box around the source in samples
                      (b1=, b2=; default br/d + 0.5 = 1) */
    int l;         /* derivative direction: 0 depth axis, 1 lateral axis (l=, default 1) */
    int nthreads;  /* OpenMP threads over positions; 0: all */
} kmig_eik_config;

void kmig_eik_config_default(kmig_eik_config *c);

/* vel: [n2][n1] velocity (n1 = depth fastest), axes o1/d1 (depth), o2/d2
   (lateral).  shots: [nshot][3] source coordinates (depth, lateral,
   cross-line), as the rows of eikods2d's shotfile= (the cross-line one is
   the model's o3, normally 0).  Outputs [nshot][n2][n1] each; tdl1, gradz
   and gradx may be null.  Returns KMIG_EIK_OK, KMIG_EIK_W_FRONT or KMIG_EIK_E_PARAM. */
int kmig_eikods2d(const kmig_eik_config *cfg, const float *vel, int n1, int n2, float o1,
                  float o2, float d1, float d2, const float *shots, int nshot, float *time,
                  float *tdl1, float *tds1, float *gradz, float *gradx);

#ifdef __cplusplus
}
#endif

#endif
