// ctypes entry points around the reference's segmented butterfly
// (SERECKIRCH/src/bf1d.cpp), compiled from that source where it lies by
// oracle/ref.mk into oracle/_ref/libbf1d_ref.so.  TEST INFRASTRUCTURE ONLY:
// tests/golden/make_bf_golden.py uses it to write the butterfly fixtures and
// the CPU tests use it, when it is built, to pin the oracle's restatement.
//
// The reference file is included as a translation unit (not copied) so that
// the panel list of a segmented factor can be reported.
#include "bf1d.cpp"

#include <vector>

namespace {

struct Mat {
    std::vector<float *> rows;
    Mat(int n, const float *tau)
    {
        rows.resize((size_t)n);
        for (int r = 0; r < n; r++) rows[(size_t)r] = const_cast<float *>(tau) + (size_t)r * n;
    }
};

}  // namespace

extern "C" {

// Segmented factor of K(r, c) = amp(r, c) exp(-i omega tau(r, c)) (tau
// row-major n x n, amp interleaved complex n x n), applied to nvec vectors
// (interleaved complex, n each).  panels (5 ints each: row0, col0, rows,
// cols, kind 0 exact / 1 butterfly) is filled up to max_panels; returns the
// panel count.
int refbf_seg_apply(int n, const float *tau, const float *amp, float omega, int p, int leaf,
                    int levels, float amp_eps, float phase_tol, int nvec, const float *uin,
                    float *uout, int *panels, int max_panels)
{
    Mat m(n, tau);
    BFStrictSegmentedFactor *F = bf1d_strict_segmented_create_phase_amp(
        n, m.rows.data(), reinterpret_cast<const fftwf_complex *>(amp), omega, p, leaf, levels,
        amp_eps, phase_tol);
    for (int v = 0; v < nvec; v++)
        bf1d_strict_segmented_apply(F, reinterpret_cast<const fftwf_complex *>(uin) + (size_t)v * n,
                                    reinterpret_cast<fftwf_complex *>(uout) + (size_t)v * n);
    const int np = (int)F->panels.size();
    for (int k = 0; k < np && k < max_panels && panels; k++) {
        const BFPanel &P = F->panels[(size_t)k];
        int *o = panels + 5 * (size_t)k;
        o[0] = P.row0;
        o[1] = P.col0;
        o[2] = P.rows;
        o[3] = P.cols;
        o[4] = P.exact_leaf ? 0 : 1;
    }
    bf1d_strict_segmented_destroy(F);
    return np;
}

// The strict (single-panel) factor, n = leaf 2^L, applied to nvec vectors.
void refbf_strict_apply(int n, const float *tau, const float *amp, float omega, int p, int leaf,
                        int nvec, const float *uin, float *uout)
{
    Mat m(n, tau);
    BFStrictFactor *F = bf1d_strict_factor_create_phase_amp(
        n, m.rows.data(), reinterpret_cast<const fftwf_complex *>(amp), omega, p, leaf);
    for (int v = 0; v < nvec; v++)
        bf1d_strict_factor_apply(F, reinterpret_cast<const fftwf_complex *>(uin) + (size_t)v * n,
                                 reinterpret_cast<fftwf_complex *>(uout) + (size_t)v * n);
    bf1d_strict_factor_destroy(F);
}

}  // extern "C"
