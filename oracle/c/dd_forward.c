/* FP64 C restatement of the oracle's DD forward projection of one view.
 *
 * TEST INFRASTRUCTURE, not product code: only the tests load it, as the checker.
 *
 * It computes what forward_project_view() of oracle/dd_oracle.py computes, with the same
 * operations in the same order (distance_project_on_{x,y,z}3, python/GPUFuncs_pycuda.py:398-587):
 * for every slab n of the view's slab axis and every detector cell, the cell's four boundary lines
 * t = a s + b at the slab centre c give a transverse footprint [t1, t2] and an axial one
 * [v1, v2]; the DD overlaps w1 of the transverse cells and w2 of the axial cells
 * (python/GPUFuncs_pycuda.py:427-452; weights below zero_tol zeroed) weigh the slab's voxels as
 * ((f w1) w2) L, summed transverse cell outer, axial cell inner, per slab, and the slab sums
 * added in slab order.  The per-cell line tables (aU1 .. bV2, L), the voxel planes and the axis
 * mapping come from the caller (the NumPy oracle's fp_view_tables, planes and AXES), so this file
 * holds only the loop.  Voxels are read in place from the float32 volume through per-axis
 * strides; cells whose footprint leaves the grid skip those voxels instead of adding 0.
 * Build: gcc -O3 -fopenmp -ffp-contract=off.  Slabs in order, OpenMP threads over the cells.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <omp.h>

static inline double fmin2(double a, double b) { return b < a ? b : a; }
static inline double fmax2(double a, double b) { return b > a ? b : a; }

/* as overlap_weight() in dd_backproject.c: cell k of planes for the footprint [c1, c2] */
static inline double fp_weight(double c1, double c2, int64_t s, int64_t e, int64_t k,
                               const double* planes, double pitch, int decreasing, double zero_tol)
{
    double w;
    if (s == e) {
        w = 1.0;
    } else {
        const double lo = fmin2(c1, c2), hi = fmax2(c1, c2), width = fabs(c1 - c2);
        if (k == s)
            w = decreasing ? (hi - planes[k + 1]) / width : (planes[k + 1] - lo) / width;
        else if (k == e)
            w = decreasing ? (planes[k] - lo) / width : (hi - planes[k]) / width;
        else
            w = fabs(pitch) / width;
    }
    if (zero_tol > 0.0 && fabs(w) < zero_tol) w = 0.0;
    return w;
}

int ddo_forward_view(const float* vol, int64_t str_s, int64_t str_t, int64_t str_v,
                     int ns, int nt, int nvv, const double* ps, const double* pt, const double* pv,
                     double ds, double dtt, double dvv, const double* aU1, const double* bU1,
                     const double* aU2, const double* bU2, const double* aV1, const double* bV1,
                     const double* aV2, const double* bV2, const double* L, int64_t ncell,
                     double zero_tol, int threads, double* out)
{
    if (!vol || !out || ns < 1 || nt < 1 || nvv < 1 || ncell < 0 || threads < 1) return 1;
    const int t_dec = dtt < 0.0, v_dec = dvv < 0.0;
    for (int64_t j = 0; j < ncell; ++j) out[j] = 0.0;
    for (int n = 0; n < ns; ++n) {
        const double c = ps[n] + ds / 2.0;
        const float* slab = vol + (size_t)n * (size_t)str_s;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t j = 0; j < ncell; ++j) {
            const double t1 = aU1[j] * c + bU1[j], t2 = aU2[j] * c + bU2[j];
            const double v1 = aV1[j] * c + bV1[j], v2 = aV2[j] * c + bV2[j];
            const int64_t it1 = (int64_t)floor((t1 - pt[0]) / dtt), it2 = (int64_t)floor((t2 - pt[0]) / dtt);
            const int64_t iv1 = (int64_t)floor((v1 - pv[0]) / dvv), iv2 = (int64_t)floor((v2 - pv[0]) / dvv);
            const int64_t st = it1 < it2 ? it1 : it2, et = it1 < it2 ? it2 : it1;
            const int64_t sv = iv1 < iv2 ? iv1 : iv2, ev = iv1 < iv2 ? iv2 : iv1;
            const int64_t k0 = st < 0 ? 0 : st, k1 = et > nt - 1 ? nt - 1 : et;
            const int64_t l0 = sv < 0 ? 0 : sv, l1 = ev > nvv - 1 ? nvv - 1 : ev;
            double acc = 0.0;
            for (int64_t k = k0; k <= k1; ++k) {
                const double w1 = fp_weight(t1, t2, st, et, k, pt, dtt, t_dec, zero_tol);
                for (int64_t l = l0; l <= l1; ++l) {
                    const double w2 = fp_weight(v1, v2, sv, ev, l, pv, dvv, v_dec, zero_tol);
                    const double f = (double)slab[k * str_t + l * str_v];
                    acc = acc + ((f * w1) * w2) * L[j];
                }
            }
            out[j] += acc;
        }
    }
    return 0;
}
