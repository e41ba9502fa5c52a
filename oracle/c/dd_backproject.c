/* FP64 C restatement of the oracle's FDK-weighted DD backprojection.
 *
 * TEST INFRASTRUCTURE, not product code: only the tests, the smoke check and the CPU-baseline
 * leg of the benchmark load it, and they use it as the checker (and, in the benchmark, as the
 * CPU time a port of the algorithm takes), never as the path being measured.
 *
 * It computes what backproject() / backproject_view() of oracle/dd_oracle.py compute, with the
 * same operations in the same order:
 *   - per view and voxel column, the eight corners of the voxel rotated into the source frame,
 *     the min/max u slope through the source (python/GPUFuncs_pycuda.py:870-897), its flat
 *     detector intercept or, for a curved detector, its angle -atan(u)
 *     (python/GPUFuncs_numba.py:580-604), and the column's weight SAD^2 / (SAD - y_rot)^2
 *     (python/GPUFuncs_pycuda.py:899-903);
 *   - per voxel, the min/max v slope over the eight corners and its detector intercept
 *     (python/GPUFuncs_pycuda.py:905-925);
 *   - the distance-driven overlap of both footprints with the detector cells, cell k taking
 *     (edge overlap)/width at the two ends, pitch/width inside and 1 when the footprint lies in
 *     one cell (python/GPUFuncs_pycuda.py:427-452, applied at :927-950);
 *   - per voxel, sum over rows k of w_v(k) * sum over columns l of q[k][l] * dtheta * w_u(l) * W,
 *     rows and columns in increasing order, added to the volume view after view
 *     (python/Reconstruction_pycuda.py:320-435).
 * Detector cell boundaries, voxel centres and per-view source heights come from the caller, so
 * the grid conventions live in one place (the NumPy oracle) and this file holds only the loop.
 *
 * Two changes of form leave every value unchanged:
 *   - the min (max) over a column's four corner denominators d_j > 0 of n / d_j is n / d_max
 *     (n / d_min) for n >= 0 and n / d_min (n / d_max) for n < 0: correctly rounded division is
 *     monotone in d, so the four divisions per z edge reduce to two (all eight are taken when a
 *     denominator is not positive);
 *   - detector columns outside [0, nu) and rows outside [0, nv) are skipped rather than added
 *     with weight 0.
 * Build: gcc -O3 -fopenmp -ffp-contract=off (no contraction, so every product and sum rounds
 * as NumPy rounds it).  OpenMP threads split the image rows (y) of each view.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#include <omp.h>

static inline double dmin2(double a, double b) { return b < a ? b : a; }
static inline double dmax2(double a, double b) { return b > a ? b : a; }

/* DD overlap weight of cell k for the footprint between intercepts c1 and c2 whose end cells
 * are s <= e; planes[k], planes[k + 1] bound cell k; `decreasing` when planes decrease
 * (the v grid, whose pitch is negative). */
static inline double overlap_weight(double c1, double c2, int64_t s, int64_t e, int64_t k,
                                    const double* planes, double pitch, int decreasing,
                                    double zero_tol)
{
    double w;
    if (s == e) {
        w = 1.0;
    } else {
        const double lo = dmin2(c1, c2), hi = dmax2(c1, c2), width = fabs(c1 - c2);
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

/* One image row's worth of per-column state for one view. */
typedef struct {
    double den[4];   /* Sy - y_r of the four corner columns (x +- hx, y +- hy) */
    double dlo, dhi; /* min / max of den */
    int positive;    /* every den > 0 */
    int64_t tap0;    /* first detector column with a weight */
    int ntap;        /* number of detector columns with a weight */
    size_t off;      /* offset of the column's weights (q-independent: w_u * W) */
} Column;

int ddo_backproject(const float* q, int n_views, int nv, int nu,
                    const double* cos_th, const double* sin_th, const double* src_z,
                    double src_x, double src_y, double det_y, double sad, double dtheta,
                    const double* u_planes, const double* v_planes, double du, double dv,
                    int curved,
                    const double* xc, const double* yc, const double* zc, int nx, int ny, int nz,
                    double hx, double hy, double hz, double zero_tol, int threads, double* vol)
{
    if (!q || !vol || n_views < 0 || nv < 1 || nu < 1 || nx < 1 || ny < 1 || nz < 1 || threads < 1)
        return 1;
    const double R = sad;
    for (int i = 0; i < n_views; ++i) {
        const float* qi = q + (size_t)i * nv * nu;
        const double c = cos_th[i], s = sin_th[i], Sz = src_z[i];
        int failed = 0;
#pragma omp parallel num_threads(threads) reduction(| : failed)
        {
            Column* col = (Column*)malloc(sizeof(Column) * (size_t)nx);
            double* wu = (double*)malloc(sizeof(double) * (size_t)nx * (size_t)(nu + 1));
            if (!col || !wu) failed = 1;
#pragma omp for schedule(dynamic, 1)
            for (int y = 0; y < ny; ++y) {
                if (failed) continue;
                const double Y = yc[y];
                size_t used = 0;
                for (int x = 0; x < nx; ++x) {
                    Column* cl = &col[x];
                    const double X = xc[x];
                    double umin = 0.0, umax = 0.0;
                    int j = 0;
                    for (int sx = 1; sx >= -1; sx -= 2) {
                        for (int sy = 1; sy >= -1; sy -= 2, ++j) {
                            const double xe = X + sx * hx, ye = Y + sy * hy;
                            const double xr = xe * c + ye * s;
                            const double yr = ye * c - xe * s;
                            const double d = src_y - yr;
                            const double u = (src_x - xr) / d;
                            cl->den[j] = d;
                            umin = j == 0 ? u : dmin2(umin, u);
                            umax = j == 0 ? u : dmax2(umax, u);
                        }
                    }
                    cl->dlo = dmin2(dmin2(cl->den[0], cl->den[1]), dmin2(cl->den[2], cl->den[3]));
                    cl->dhi = dmax2(dmax2(cl->den[0], cl->den[1]), dmax2(cl->den[2], cl->den[3]));
                    cl->positive = cl->dlo > 0.0;
                    double cu1, cu2;
                    if (curved) {
                        cu1 = -atan(umin);
                        cu2 = -atan(umax);
                    } else {
                        cu1 = umin * det_y + (src_x - umin * src_y);
                        cu2 = umax * det_y + (src_x - umax * src_y);
                    }
                    const int64_t iu1 = (int64_t)floor((cu1 - u_planes[0]) / du);
                    const int64_t iu2 = (int64_t)floor((cu2 - u_planes[0]) / du);
                    const int64_t su = iu1 < iu2 ? iu1 : iu2, eu = iu1 < iu2 ? iu2 : iu1;
                    const double y_rot = Y * c - X * s;
                    const double W = (R * R) / ((R - y_rot) * (R - y_rot));
                    const int64_t l0 = su < 0 ? 0 : su, l1 = eu > nu - 1 ? nu - 1 : eu;
                    cl->tap0 = l0;
                    cl->ntap = l1 >= l0 ? (int)(l1 - l0 + 1) : 0;
                    cl->off = used;
                    for (int64_t l = l0; l <= l1; ++l)
                        wu[used++] = overlap_weight(cu1, cu2, su, eu, l, u_planes, du, 0, zero_tol) * W;
                }
                for (int z = 0; z < nz; ++z) {
                    const double Z = zc[z];
                    const double n1 = Sz - (Z + hz), n2 = Sz - (Z - hz);
                    double* vrow = vol + ((size_t)z * ny + y) * nx;
                    for (int x = 0; x < nx; ++x) {
                        const Column* cl = &col[x];
                        if (cl->ntap == 0) continue;
                        double vmin, vmax;
                        if (cl->positive) {
                            const double a1 = n1 / cl->dlo, b1 = n1 / cl->dhi;
                            const double a2 = n2 / cl->dlo, b2 = n2 / cl->dhi;
                            vmin = dmin2(dmin2(a1, b1), dmin2(a2, b2));
                            vmax = dmax2(dmax2(a1, b1), dmax2(a2, b2));
                        } else {
                            vmin = vmax = n1 / cl->den[0];
                            for (int k = 0; k < 4; ++k) {
                                const double e1 = n1 / cl->den[k], e2 = n2 / cl->den[k];
                                vmin = dmin2(vmin, dmin2(e1, e2));
                                vmax = dmax2(vmax, dmax2(e1, e2));
                            }
                        }
                        const double cv1 = vmin * det_y + (Sz - vmin * src_y);
                        const double cv2 = vmax * det_y + (Sz - vmax * src_y);
                        const int64_t iv1 = (int64_t)floor((cv1 - v_planes[0]) / dv);
                        const int64_t iv2 = (int64_t)floor((cv2 - v_planes[0]) / dv);
                        const int64_t sv = iv1 < iv2 ? iv1 : iv2, ev = iv1 < iv2 ? iv2 : iv1;
                        const int64_t k0 = sv < 0 ? 0 : sv, k1 = ev > nv - 1 ? nv - 1 : ev;
                        if (k1 < k0) continue;
                        const double* w = wu + cl->off;
                        double acc = 0.0;
                        for (int64_t k = k0; k <= k1; ++k) {
                            const double w1 = overlap_weight(cv1, cv2, sv, ev, k, v_planes, dv, 1, zero_tol);
                            const float* qrow = qi + (size_t)k * nu + cl->tap0;
                            double row = 0.0;
                            for (int t = 0; t < cl->ntap; ++t)
                                row = row + ((double)qrow[t] * dtheta) * w[t];
                            acc = acc + row * w1;
                        }
                        vrow[x] += acc;
                    }
                }
            }
            free(col);
            free(wu);
        }
        if (failed) return 2;
    }
    return 0;
}
