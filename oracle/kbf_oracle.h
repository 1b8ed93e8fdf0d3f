/* CPU oracle for the segmented butterfly of kirchdat2d_freq (use_bf=1).
   TEST INFRASTRUCTURE ONLY. */
#ifndef KBF_ORACLE_H
#define KBF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int p;           /* bf_p: Chebyshev nodes per box */
    int leaf;        /* bf_n_leaf: leaf box size */
    int levels;      /* bf_panel_levels: top panels of leaf * 2^levels */
    float amp_eps;   /* bf_amp_eps: |amp| at or below it counts as zero */
    float phase_tol; /* bf_phase_tol: largest non-separable phase of a butterfly panel */
} kbo_cfg;

/* Panel kinds reported by kbo_seg_apply. */
#define KBO_EXACT 0
#define KBO_BUTTERFLY 1

/* The segmented operator of K(r, c) = amp(r, c) exp(-i omega tau(r, c)) (tau
   row-major n x n, amp interleaved complex n x n) applied to nvec vectors
   (interleaved complex, n each).  panels (5 ints each: row0, col0, rows,
   cols, kind) is filled up to max_panels when not NULL.  Returns the panel
   count, or -1 for arguments the reference rejects (n <= 0, p <= 0,
   leaf <= 1, p >= leaf). */
int kbo_seg_apply(int n, const float *tau, const float *amp, float omega, const kbo_cfg *cfg,
                  int nvec, const float *uin, float *uout, int *panels, int max_panels);

/* One strict butterfly factor (n = leaf 2^L, every entry active) applied to
   nvec vectors.  Returns 0, or -1 when n is not leaf times a power of two or
   p >= leaf. */
int kbo_strict_apply(int n, const float *tau, const float *amp, float omega, int p, int leaf,
                     int nvec, const float *uin, float *uout);

/* Lagrange basis of the p nodes at x (double), as the factor uses it. */
void kbo_lagrange(int p, const double *nodes, double x, double *basis);

#ifdef __cplusplus
}
#endif
#endif
