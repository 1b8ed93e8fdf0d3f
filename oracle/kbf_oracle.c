/*
 * CPU oracle for the segmented butterfly that kirchdat2d_freq applies per
 * frequency when use_bf=1 (SERECKIRCH/src/bf1d.cpp).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ (and through kmdat_oracle.c by
 * the datuming oracle), never by the product library.
 *
 * The operator is K(r, c) = amp(r, c) exp(-i omega tau(r, c)) on an n x n
 * matrix (r: output row, c: input column).  Restated (file:line in
 * SERECKIRCH/src/bf1d.cpp):
 *   :61-96     boxes of a level (2^l boxes of n / 2^l points; centre, width
 *              i2 - i1, p Chebyshev points of the first kind across it)
 *   :98-167    barycentric Lagrange weights and basis (Berrut & Trefethen,
 *              SIAM Review 46 (2004), the second barycentric form), with the
 *              exact-node shortcut
 *   :169-240   bilinear interpolation of tau and of the complex amplitude at
 *              real (column, row) coordinates, with the exact-sample shortcut
 *   :283-473   the strict factor's precomputed stages: leaf initialisation,
 *              source-side recursion to the middle level, the switch kernel,
 *              target-side recursion, termination
 *   :522-645   the strict factor applied to one vector
 *   :703-760   panel statistics and exact leaf panels (K = amp exp(-i omega
 *              tau) in float, entries with |amp|^2 <= amp_eps^2 dropped)
 *   :816-933   the non-separable phase estimate on a 5 x 5 sample grid, the
 *              legality of a butterfly panel and the adaptive split
 *   :935-1033  top panels of leaf 2^levels, the segmented apply
 * Arithmetic follows the reference's: positions, Lagrange values and
 * interpolated traveltimes in double, the phase (float)(alpha t) through
 * float cos/sin, every product and sum of the apply in float.
 *
 * How it is pinned (tests/test_bf_oracle.py): against fixtures written by
 * the reference's own bf1d.cpp compiled from source (tests/golden/bf.npz);
 * against the compiled reference directly when oracle/_ref holds it; and by
 * closed forms: with omega = 0 and a bilinear amplitude every stage
 * interpolates exactly, and a separable phase with a constant amplitude is a
 * rank-one kernel the factor reproduces, both to float rounding; the
 * Lagrange basis reproduces polynomials of degree < p.
 */
#include "kbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    float re, im;
} cf;

static cf cf_mul(cf a, cf b)
{
    cf z;
    z.re = a.re * b.re - a.im * b.im;
    z.im = a.re * b.im + a.im * b.re;
    return z;
}

static cf cf_scale(cf a, float s)
{
    cf z;
    z.re = a.re * s;
    z.im = a.im * s;
    return z;
}

static void cf_add(cf *a, cf b)
{
    a->re += b.re;
    a->im += b.im;
}

/* exp(i ph) with the phase rounded to float first */
static cf cf_phase(double ph)
{
    const float f = (float)ph;
    cf z;
    z.re = cosf(f);
    z.im = sinf(f);
    return z;
}

/* ------------------------------------------------------------- geometry -- */

/* The boxes of one block of n = leaf 2^L points: level l has 2^l boxes, box b
   of level l is entry (2^l - 1) + b. */
typedef struct {
    int L, p, nbox;
    int *i1;
    double *x0;   /* centre */
    double *node; /* [box][p] */
} Tree;

static int tree_init(Tree *t, int n, int leaf, int p)
{
    int L = 0;
    while ((leaf << L) < n) L++;
    t->L = L;
    t->p = p;
    t->nbox = (2 << L) - 1;
    t->i1 = (int *)malloc(sizeof(int) * t->nbox);
    t->x0 = (double *)malloc(sizeof(double) * t->nbox);
    t->node = (double *)malloc(sizeof(double) * t->nbox * p);
    double *cheb = (double *)malloc(sizeof(double) * p);
    if (!t->i1 || !t->x0 || !t->node || !cheb) {
        free(cheb);
        return -1;
    }
    for (int k = 0; k < p; k++) cheb[k] = cos(((2.0 * (double)(k + 1) - 1.0) * M_PI) / (2.0 * (double)p));
    for (int l = 0; l <= L; l++) {
        const int nb = 1 << l, size = n / nb;
        for (int b = 0; b < nb; b++) {
            const int e = nb - 1 + b;
            const double lo = (double)(b * size), hi = (double)((b + 1) * size - 1);
            const double c = 0.5 * (lo + hi), width = hi - lo;
            t->i1[e] = b * size;
            t->x0[e] = c;
            for (int k = 0; k < p; k++) t->node[(size_t)e * p + k] = c + 0.5 * width * cheb[k];
        }
    }
    free(cheb);
    return 0;
}

static void tree_free(Tree *t)
{
    free(t->i1);
    free(t->x0);
    free(t->node);
}

static int box_of(int l, int b) { return (1 << l) - 1 + b; }

void kbo_lagrange(int p, const double *nodes, double x, double *basis)
{
    for (int j = 0; j < p; j++) basis[j] = 0.0;
    for (int j = 0; j < p; j++)
        if (fabs(x - nodes[j]) < 1e-13) {
            basis[j] = 1.0;
            return;
        }
    double sum = 0.0;
    for (int j = 0; j < p; j++) {
        double wj = 1.0;
        for (int k = 0; k < p; k++)
            if (k != j) wj /= (nodes[j] - nodes[k]);
        basis[j] = wj / (x - nodes[j]);
        sum += basis[j];
    }
    for (int j = 0; j < p; j++) basis[j] /= sum;
}

/* Bilinear sample of a row-major n x n float matrix at column x, row z. */
static double bilinear(const float *A, int n, double x, double z)
{
    if (x < 0.0) x = 0.0;
    if (x > (double)(n - 1)) x = (double)(n - 1);
    if (z < 0.0) z = 0.0;
    if (z > (double)(n - 1)) z = (double)(n - 1);
    const long jr = lround(x), ir = lround(z);
    if (fabs(x - (double)jr) < 1e-12 && fabs(z - (double)ir) < 1e-12)
        return (double)A[(size_t)ir * n + jr];
    int j = (int)floor(x), i = (int)floor(z);
    if (j > n - 2) j = n - 2;
    if (i > n - 2) i = n - 2;
    if (j < 0) j = 0;
    if (i < 0) i = 0;
    const double fx = x - (double)j, fz = z - (double)i;
    const float *r0 = A + (size_t)i * n, *r1 = r0 + n;
    return (1.0 - fx) * (1.0 - fz) * (double)r0[j] + fx * (1.0 - fz) * (double)r0[j + 1] +
           (1.0 - fx) * fz * (double)r1[j] + fx * fz * (double)r1[j + 1];
}

/* The same for an interleaved complex matrix, rounded to float. */
static cf bilinear_c(const float *A, int n, double x, double z)
{
    if (x < 0.0) x = 0.0;
    if (x > (double)(n - 1)) x = (double)(n - 1);
    if (z < 0.0) z = 0.0;
    if (z > (double)(n - 1)) z = (double)(n - 1);
    const long jr = lround(x), ir = lround(z);
    cf v;
    if (fabs(x - (double)jr) < 1e-12 && fabs(z - (double)ir) < 1e-12) {
        v.re = A[2 * ((size_t)ir * n + jr)];
        v.im = A[2 * ((size_t)ir * n + jr) + 1];
        return v;
    }
    int j = (int)floor(x), i = (int)floor(z);
    if (j > n - 2) j = n - 2;
    if (i > n - 2) i = n - 2;
    if (j < 0) j = 0;
    if (i < 0) i = 0;
    const double fx = x - (double)j, fz = z - (double)i;
    const double w00 = (1.0 - fx) * (1.0 - fz), w01 = fx * (1.0 - fz);
    const double w10 = (1.0 - fx) * fz, w11 = fx * fz;
    const float *a00 = A + 2 * ((size_t)i * n + j), *a01 = a00 + 2;
    const float *a10 = a00 + 2 * (size_t)n, *a11 = a10 + 2;
    v.re = (float)(w00 * a00[0] + w01 * a01[0] + w10 * a10[0] + w11 * a11[0]);
    v.im = (float)(w00 * a00[1] + w01 * a01[1] + w10 * a10[1] + w11 * a11[1]);
    return v;
}

/* ------------------------------------------------------- strict factor -- */

/* Lagrange values of `npt` points against the p nodes of one box, rounded
   to float as the apply uses them: out[pt][node]. */
static void lagrange_table(int p, const double *nodes, const double *pts, int npt, float *out)
{
    double *b = (double *)malloc(sizeof(double) * p);
    for (int k = 0; k < npt; k++) {
        kbo_lagrange(p, nodes, pts[k], b);
        for (int j = 0; j < p; j++) out[(size_t)k * p + j] = (float)b[j];
    }
    free(b);
}

/* One strict factor: an n x n block (n = leaf 2^L, every entry active)
   with its stage tables formed once.  The interpolation coefficients of a
   level are indexed [source box][target box][node]; stage lv of the source
   recursion has 2^lv target boxes and 2^(L-lv) source boxes, the switch sits
   at level mid = L / 2, the target recursion runs from mid + 1 to L. */
typedef struct {
    float *L;               /* Lagrange values, [parent][child][child node][parent node] */
    cf *ph_in, *ph_out;     /* phases of the children's nodes and of the parents' */
} Stage;

typedef struct {
    int n, p, leaf, L, mid;
    Tree T;
    float *in_L;            /* [leaf box][point][node] */
    cf *in_pt, *in_node;    /* [point], [leaf box][node] */
    Stage *src;             /* levels 1 .. mid */
    cf *sw;                 /* [source box][target box][target node][source node] */
    Stage *tgt;             /* levels mid + 1 .. L */
    float *out_L;           /* [leaf box][point][node] */
    cf *out_node, *out_pt;
} Strict;

static double tq(const float *tau, int n, double col, double row) { return bilinear(tau, n, col, row); }

static int strict_build(Strict *S, int n, const float *tau, const float *amp, float omega, int p,
                        int leaf)
{
    memset(S, 0, sizeof *S);
    if (tree_init(&S->T, n, leaf, p)) return -1;
    const Tree *T = &S->T;
    const int L = T->L, mid = L / 2, nleaf = 1 << L;
    const double a = -(double)omega;
    S->n = n;
    S->p = p;
    S->leaf = leaf;
    S->L = L;
    S->mid = mid;
    double *pts = (double *)malloc(sizeof(double) * (size_t)leaf);

    /* leaf initialisation: points of each source leaf box against its nodes,
       phases towards the target root's centre */
    const double rc = T->x0[0];
    S->in_L = (float *)malloc(sizeof(float) * (size_t)nleaf * leaf * p);
    S->in_pt = (cf *)malloc(sizeof(cf) * (size_t)n);
    S->in_node = (cf *)malloc(sizeof(cf) * (size_t)nleaf * p);
    for (int bs = 0; bs < nleaf; bs++) {
        const int e = box_of(L, bs);
        const double *nd = T->node + (size_t)e * p;
        for (int k = 0; k < leaf; k++) {
            pts[k] = (double)(T->i1[e] + k);
            S->in_pt[T->i1[e] + k] = cf_phase(a * tq(tau, n, pts[k], rc));
        }
        lagrange_table(p, nd, pts, leaf, S->in_L + (size_t)bs * leaf * p);
        for (int it = 0; it < p; it++)
            S->in_node[(size_t)bs * p + it] = cf_phase(-a * tq(tau, n, nd[it], rc));
    }

    /* source-side recursion */
    S->src = (Stage *)calloc((size_t)mid + 1, sizeof(Stage));
    for (int lv = 1; lv <= mid; lv++) {
        Stage *G = &S->src[lv];
        const int nr = 1 << lv, ns_ = 1 << (L - lv), nsc = 2 * ns_;
        G->L = (float *)malloc(sizeof(float) * (size_t)nsc * p * p);
        G->ph_in = (cf *)malloc(sizeof(cf) * (size_t)nr * nsc * p);
        G->ph_out = (cf *)malloc(sizeof(cf) * (size_t)nr * ns_ * p);
        for (int bs = 0; bs < ns_; bs++)
            for (int ch = 0; ch < 2; ch++)
                lagrange_table(p, T->node + (size_t)box_of(L - lv, bs) * p,
                               T->node + (size_t)box_of(L - lv + 1, 2 * bs + ch) * p, p,
                               G->L + (size_t)(2 * bs + ch) * p * p);
        for (int br = 0; br < nr; br++) {
            const double r0 = T->x0[box_of(lv, br)];
            for (int cs = 0; cs < nsc; cs++)
                for (int jt = 0; jt < p; jt++)
                    G->ph_in[((size_t)br * nsc + cs) * p + jt] = cf_phase(
                        a * tq(tau, n, T->node[(size_t)box_of(L - lv + 1, cs) * p + jt], r0));
            for (int bs = 0; bs < ns_; bs++)
                for (int it = 0; it < p; it++)
                    G->ph_out[((size_t)br * ns_ + bs) * p + it] = cf_phase(
                        -a * tq(tau, n, T->node[(size_t)box_of(L - lv, bs) * p + it], r0));
        }
    }

    /* switch kernel at the middle level: amplitude and traveltime
       interpolated bilinearly between source and target nodes */
    {
        const int nr = 1 << mid, ns_ = 1 << (L - mid);
        S->sw = (cf *)malloc(sizeof(cf) * (size_t)ns_ * nr * p * p);
        for (int bs = 0; bs < ns_; bs++) {
            const double *sn = T->node + (size_t)box_of(L - mid, bs) * p;
            for (int br = 0; br < nr; br++) {
                const double *tn = T->node + (size_t)box_of(mid, br) * p;
                for (int it = 0; it < p; it++)
                    for (int jt = 0; jt < p; jt++)
                        S->sw[(((size_t)bs * nr + br) * p + it) * p + jt] =
                            cf_mul(bilinear_c(amp, n, sn[jt], tn[it]),
                                   cf_phase(a * tq(tau, n, sn[jt], tn[it])));
            }
        }
    }

    /* target-side recursion: each target box interpolates from its parent's
       nodes; phases from the centres of the source children */
    S->tgt = (Stage *)calloc((size_t)L + 1, sizeof(Stage));
    for (int lv = mid + 1; lv <= L; lv++) {
        Stage *G = &S->tgt[lv];
        const int nr = 1 << lv, nsc = 1 << (L - lv + 1);
        G->L = (float *)malloc(sizeof(float) * (size_t)nr * p * p);
        G->ph_in = (cf *)malloc(sizeof(cf) * (size_t)nr * nsc * p);
        G->ph_out = (cf *)malloc(sizeof(cf) * (size_t)nr * nsc * p);
        for (int br = 0; br < nr; br++) {
            const int et = box_of(lv, br), ep = box_of(lv - 1, br / 2);
            lagrange_table(p, T->node + (size_t)ep * p, T->node + (size_t)et * p, p,
                           G->L + (size_t)br * p * p);
            for (int cs = 0; cs < nsc; cs++) {
                const double s0 = T->x0[box_of(L - lv + 1, cs)];
                for (int jt = 0; jt < p; jt++)
                    G->ph_in[((size_t)br * nsc + cs) * p + jt] =
                        cf_phase(-a * tq(tau, n, s0, T->node[(size_t)ep * p + jt]));
                for (int it = 0; it < p; it++)
                    G->ph_out[((size_t)br * nsc + cs) * p + it] =
                        cf_phase(a * tq(tau, n, s0, T->node[(size_t)et * p + it]));
            }
        }
    }

    /* termination: target leaf nodes to the rows, phases from the source
       root's centre */
    const double sc = T->x0[0];
    S->out_L = (float *)malloc(sizeof(float) * (size_t)nleaf * leaf * p);
    S->out_node = (cf *)malloc(sizeof(cf) * (size_t)nleaf * p);
    S->out_pt = (cf *)malloc(sizeof(cf) * (size_t)n);
    for (int br = 0; br < nleaf; br++) {
        const int e = box_of(L, br);
        const double *nd = T->node + (size_t)e * p;
        for (int k = 0; k < leaf; k++) {
            pts[k] = (double)(T->i1[e] + k);
            S->out_pt[T->i1[e] + k] = cf_phase(a * tq(tau, n, sc, pts[k]));
        }
        lagrange_table(p, nd, pts, leaf, S->out_L + (size_t)br * leaf * p);
        for (int it = 0; it < p; it++)
            S->out_node[(size_t)br * p + it] = cf_phase(-a * tq(tau, n, sc, nd[it]));
    }
    free(pts);
    return 0;
}

static void strict_free(Strict *S)
{
    for (int lv = 1; lv <= S->mid; lv++) {
        free(S->src[lv].L);
        free(S->src[lv].ph_in);
        free(S->src[lv].ph_out);
    }
    for (int lv = S->mid + 1; lv <= S->L; lv++) {
        free(S->tgt[lv].L);
        free(S->tgt[lv].ph_in);
        free(S->tgt[lv].ph_out);
    }
    free(S->src);
    free(S->tgt);
    free(S->in_L);
    free(S->in_pt);
    free(S->in_node);
    free(S->sw);
    free(S->out_L);
    free(S->out_node);
    free(S->out_pt);
    tree_free(&S->T);
}

/* One vector through the factor; sa, sb hold (n / leaf) p coefficients. */
static void strict_apply1(const Strict *S, const cf *u, cf *out, cf *sa, cf *sb)
{
    const int p = S->p, leaf = S->leaf, L = S->L, mid = S->mid, nleaf = 1 << L;
    const Tree *T = &S->T;
    cf acc[128], tmp[128];
    cf *prev = sa, *cur = sb;

    for (int bs = 0; bs < nleaf; bs++) {
        const int i1 = T->i1[box_of(L, bs)];
        const float *Lg = S->in_L + (size_t)bs * leaf * p;
        for (int it = 0; it < p; it++) {
            cf s = {0.0f, 0.0f};
            for (int k = 0; k < leaf; k++)
                cf_add(&s, cf_scale(cf_mul(u[i1 + k], S->in_pt[i1 + k]), Lg[(size_t)k * p + it]));
            prev[(size_t)bs * p + it] = cf_mul(s, S->in_node[(size_t)bs * p + it]);
        }
    }
    for (int lv = 1; lv <= mid; lv++) {
        const Stage *G = &S->src[lv];
        const int nr = 1 << lv, ns_ = 1 << (L - lv), nsc = 2 * ns_, nr_prev = nr / 2;
        for (int br = 0; br < nr; br++)
            for (int bs = 0; bs < ns_; bs++) {
                for (int it = 0; it < p; it++) acc[it].re = acc[it].im = 0.0f;
                for (int ch = 0; ch < 2; ch++) {
                    const int cs = 2 * bs + ch;
                    const float *Lc = G->L + (size_t)cs * p * p;
                    for (int jt = 0; jt < p; jt++) {
                        const cf w = cf_mul(prev[((size_t)cs * nr_prev + br / 2) * p + jt],
                                            G->ph_in[((size_t)br * nsc + cs) * p + jt]);
                        for (int it = 0; it < p; it++) cf_add(&acc[it], cf_scale(w, Lc[(size_t)jt * p + it]));
                    }
                }
                for (int it = 0; it < p; it++)
                    cur[((size_t)bs * nr + br) * p + it] =
                        cf_mul(acc[it], G->ph_out[((size_t)br * ns_ + bs) * p + it]);
            }
        cf *t = prev;
        prev = cur;
        cur = t;
    }
    {
        const int nr = 1 << mid, ns_ = 1 << (L - mid);
        for (int bs = 0; bs < ns_; bs++)
            for (int br = 0; br < nr; br++) {
                const cf *k = S->sw + ((size_t)bs * nr + br) * p * p;
                const cf *v = prev + ((size_t)bs * nr + br) * p;
                for (int it = 0; it < p; it++) {
                    cf s = {0.0f, 0.0f};
                    for (int jt = 0; jt < p; jt++) cf_add(&s, cf_mul(k[(size_t)it * p + jt], v[jt]));
                    cur[((size_t)bs * nr + br) * p + it] = s;
                }
            }
        cf *t = prev;
        prev = cur;
        cur = t;
    }
    for (int lv = mid + 1; lv <= L; lv++) {
        const Stage *G = &S->tgt[lv];
        const int nr = 1 << lv, ns_ = 1 << (L - lv), nsc = 2 * ns_, nr_prev = nr / 2;
        for (int br = 0; br < nr; br++) {
            const float *Lt = G->L + (size_t)br * p * p;
            for (int bs = 0; bs < ns_; bs++) {
                for (int it = 0; it < p; it++) acc[it].re = acc[it].im = 0.0f;
                for (int ch = 0; ch < 2; ch++) {
                    const int cs = 2 * bs + ch;
                    for (int jt = 0; jt < p; jt++)
                        tmp[jt] = cf_mul(prev[((size_t)cs * nr_prev + br / 2) * p + jt],
                                         G->ph_in[((size_t)br * nsc + cs) * p + jt]);
                    for (int it = 0; it < p; it++) {
                        cf y = {0.0f, 0.0f};
                        for (int jt = 0; jt < p; jt++) cf_add(&y, cf_scale(tmp[jt], Lt[(size_t)it * p + jt]));
                        cf_add(&acc[it], cf_mul(y, G->ph_out[((size_t)br * nsc + cs) * p + it]));
                    }
                }
                for (int it = 0; it < p; it++) cur[((size_t)bs * nr + br) * p + it] = acc[it];
            }
        }
        cf *t = prev;
        prev = cur;
        cur = t;
    }
    for (int br = 0; br < nleaf; br++) {
        const int i1 = T->i1[box_of(L, br)];
        const float *Lg = S->out_L + (size_t)br * leaf * p;
        for (int it = 0; it < p; it++)
            tmp[it] = cf_mul(prev[(size_t)br * p + it], S->out_node[(size_t)br * p + it]);
        for (int k = 0; k < leaf; k++) {
            cf s = {0.0f, 0.0f};
            for (int it = 0; it < p; it++) cf_add(&s, cf_scale(tmp[it], Lg[(size_t)k * p + it]));
            out[i1 + k] = cf_mul(s, S->out_pt[i1 + k]);
        }
    }
}

static int is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

static int strict_run(int n, const float *tau, const float *amp, float omega, int p, int leaf,
                      int nvec, const cf *uin, size_t in_stride, cf *uout, size_t out_stride)
{
    if (p <= 0 || p > 128 || leaf <= 1 || p >= leaf || n < leaf || n % leaf || !is_pow2(n / leaf))
        return -1;
    Strict S;
    if (strict_build(&S, n, tau, amp, omega, p, leaf)) return -1;
    const size_t ns = (size_t)(n / leaf) * p;
    cf *sa = (cf *)malloc(sizeof(cf) * ns), *sb = (cf *)malloc(sizeof(cf) * ns);
    for (int v = 0; v < nvec; v++)
        strict_apply1(&S, uin + (size_t)v * in_stride, uout + (size_t)v * out_stride, sa, sb);
    free(sa);
    free(sb);
    strict_free(&S);
    return 0;
}

int kbo_strict_apply(int n, const float *tau, const float *amp, float omega, int p, int leaf,
                     int nvec, const float *uin, float *uout)
{
    return strict_run(n, tau, amp, omega, p, leaf, nvec, (const cf *)uin, (size_t)n, (cf *)uout,
                      (size_t)n);
}

/* ------------------------------------------------------ segmented form -- */

typedef struct {
    int n;
    const float *tau, *amp;
    float omega, eps2, tol;
    int p, leaf;
    int *panels, max_panels, npanel;
    int nvec;
    const cf *uin;
    cf *uout;
} Seg;

static int active(const Seg *G, int r, int c)
{
    const float *a = G->amp + 2 * ((size_t)r * G->n + c);
    return a[0] * a[0] + a[1] * a[1] > G->eps2;
}

/* largest |omega (tau(r, c) - tau(r, c1) - tau(r1, c) + tau(r1, c1))| over
   the panel's first, last, middle and quarter rows and columns */
static double residual(const Seg *G, int r0, int c0, int rows, int cols)
{
    if (rows <= 1 || cols <= 1) return 0.0;
    const int n = G->n, r1 = r0, r2 = r0 + rows - 1, c1 = c0, c2 = c0 + cols - 1;
    const int rs[5] = {r1, r2, r0 + rows / 2, r0 + rows / 4, r0 + (3 * rows) / 4};
    const int cs[5] = {c1, c2, c0 + cols / 2, c0 + cols / 4, c0 + (3 * cols) / 4};
#define TAU(r_, c_) ((double)G->tau[(size_t)(r_) * n + (c_)])
    const double t00 = TAU(r1, c1);
    double m = 0.0;
    for (int a = 0; a < 5; a++) {
        const int rr = rs[a] < r1 ? r1 : (rs[a] > r2 ? r2 : rs[a]);
        for (int b = 0; b < 5; b++) {
            const int cc = cs[b] < c1 ? c1 : (cs[b] > c2 ? c2 : cs[b]);
            const double sep = TAU(rr, c1) + TAU(r1, cc) - t00;
            const double d = fabs((double)G->omega * (TAU(rr, cc) - sep));
            if (d > m) m = d;
        }
    }
#undef TAU
    return m;
}

static void record(Seg *G, int r0, int c0, int rows, int cols, int kind)
{
    if (G->panels && G->npanel < G->max_panels) {
        int *o = G->panels + 5 * (size_t)G->npanel;
        o[0] = r0;
        o[1] = c0;
        o[2] = rows;
        o[3] = cols;
        o[4] = kind;
    }
    G->npanel++;
}

static void exact_panel(Seg *G, int r0, int c0, int rows, int cols)
{
    const int n = G->n;
    record(G, r0, c0, rows, cols, KBO_EXACT);
    for (int v = 0; v < G->nvec; v++) {
        const cf *u = G->uin + (size_t)v * n;
        cf *y = G->uout + (size_t)v * n;
        for (int r = 0; r < rows; r++) {
            cf s = {0.0f, 0.0f};
            for (int c = 0; c < cols; c++) {
                if (!active(G, r0 + r, c0 + c)) continue;
                const size_t e = (size_t)(r0 + r) * n + c0 + c;
                const float ar = G->amp[2 * e], ai = G->amp[2 * e + 1];
                const float ph = G->omega * G->tau[e];
                const float cp = cosf(ph), sp = sinf(ph);
                cf k;
                k.re = ar * cp + ai * sp;
                k.im = ai * cp - ar * sp;
                cf_add(&s, cf_mul(k, u[c0 + c]));
            }
            y[r0 + r].re += s.re;
            y[r0 + r].im += s.im;
        }
    }
}

static void butterfly_panel(Seg *G, int r0, int c0, int m)
{
    const int n = G->n;
    record(G, r0, c0, m, m, KBO_BUTTERFLY);
    float *tl = (float *)malloc(sizeof(float) * (size_t)m * m);
    float *al = (float *)malloc(sizeof(float) * 2 * (size_t)m * m);
    cf *ui = (cf *)malloc(sizeof(cf) * (size_t)m * G->nvec);
    cf *uo = (cf *)malloc(sizeof(cf) * (size_t)m * G->nvec);
    for (int r = 0; r < m; r++)
        for (int c = 0; c < m; c++) {
            const size_t e = (size_t)(r0 + r) * n + c0 + c, l = (size_t)r * m + c;
            tl[l] = G->tau[e];
            al[2 * l] = G->amp[2 * e];
            al[2 * l + 1] = G->amp[2 * e + 1];
        }
    for (int v = 0; v < G->nvec; v++)
        memcpy(ui + (size_t)v * m, G->uin + (size_t)v * n + c0, sizeof(cf) * m);
    strict_run(m, tl, al, G->omega, G->p, G->leaf, G->nvec, ui, (size_t)m, uo, (size_t)m);
    for (int v = 0; v < G->nvec; v++)
        for (int r = 0; r < m; r++) {
            cf *y = G->uout + (size_t)v * n + r0 + r;
            y->re += uo[(size_t)v * m + r].re;
            y->im += uo[(size_t)v * m + r].im;
        }
    free(tl);
    free(al);
    free(ui);
    free(uo);
}

static void adaptive(Seg *G, int r0, int c0, int rows, int cols)
{
    if (rows <= 0 || cols <= 0) return;
    long nnz = 0;
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) nnz += active(G, r0 + r, c0 + c);
    if (nnz == 0) return;
    const int leaf = G->leaf;
    const int legal = rows == cols && rows >= 2 * leaf && rows % leaf == 0 && is_pow2(rows / leaf);
    if (nnz == (long)rows * cols && legal && residual(G, r0, c0, rows, cols) <= (double)G->tol) {
        butterfly_panel(G, r0, c0, rows);
        return;
    }
    if (rows <= leaf && cols <= leaf) {
        exact_panel(G, r0, c0, rows, cols);
    } else if (rows >= cols && rows > leaf) {
        const int h = rows / 2;
        adaptive(G, r0, c0, h, cols);
        adaptive(G, r0 + h, c0, rows - h, cols);
    } else if (cols > leaf) {
        const int h = cols / 2;
        adaptive(G, r0, c0, rows, h);
        adaptive(G, r0, c0 + h, rows, cols - h);
    } else {
        exact_panel(G, r0, c0, rows, cols);
    }
}

int kbo_seg_apply(int n, const float *tau, const float *amp, float omega, const kbo_cfg *cfg,
                  int nvec, const float *uin, float *uout, int *panels, int max_panels)
{
    if (!cfg || n <= 0 || cfg->p <= 0 || cfg->leaf <= 1 || cfg->p >= cfg->leaf) return -1;
    Seg G;
    G.n = n;
    G.tau = tau;
    G.amp = amp;
    G.omega = omega;
    const float eps = cfg->amp_eps < 0.0f ? 0.0f : cfg->amp_eps;
    G.eps2 = eps * eps;
    G.tol = (cfg->phase_tol <= 0.0f || !isfinite(cfg->phase_tol)) ? 2.0f : cfg->phase_tol;
    G.p = cfg->p;
    G.leaf = cfg->leaf;
    G.panels = panels;
    G.max_panels = max_panels;
    G.npanel = 0;
    G.nvec = nvec;
    G.uin = (const cf *)uin;
    G.uout = (cf *)uout;
    memset(uout, 0, sizeof(cf) * (size_t)n * nvec);
    /* top panels of leaf 2^levels (levels < 0 reads as 2), at least a leaf,
       at most the smallest leaf 2^k covering n */
    const int lev = cfg->levels < 0 ? 2 : cfg->levels;
    int pn = cfg->leaf;
    for (int k = 0; k < lev; k++) pn <<= 1;
    if (pn < cfg->leaf) pn = cfg->leaf;
    if (pn > n) {
        int m = cfg->leaf;
        while (m < n) m <<= 1;
        pn = m;
    }
    for (int r0 = 0; r0 < n; r0 += pn)
        for (int c0 = 0; c0 < n; c0 += pn)
            adaptive(&G, r0, c0, pn < n - r0 ? pn : n - r0, pn < n - c0 ? pn : n - c0);
    return G.npanel;
}
