/* FP64 C restatement of the pixel oracle's voxel-driven backprojection.
 *
 * TEST INFRASTRUCTURE, not product code: only the tests and the CPU-baseline leg of the
 * benchmark load it, as the checker and as the CPU time of a port, never as the path measured.
 *
 * It computes what backproject() and edge_margin() of oracle/pixel_oracle.py compute, with the
 * same operations in the same order:
 *   - per view and voxel column, t = x cos a + y sin a, s = -x sin a + y cos a, the detector
 *     abscissa U = R t / (R - s) and the weight W = R^2 / (R - s)^2; per voxel V = R z / (R - s)
 *     (CalculateInterpolGrid, python/ConeBeamGPU.py:116-129, used at :303-309);
 *   - the bilinear sample of Interpol2dLineargpu (python/ConeBeamGPU.py:80-102): the fractional
 *     cell indices, floor and ceil as the two nodes, node distances as weights (the first one
 *     set to 1 when both are under 1e-4, a sample on a node), 0 when a node falls off the grid;
 *   - vol += sample * scale * W view after view (python/ConeBeamGPU.py:303-315);
 *   - optionally, per voxel, the smallest distance over the views of the fractional indices
 *     from the lines where the bounds test flips (0 and n - 1 in u and in v).
 * Cell centres, voxel coordinates and view angles come from the caller (the NumPy oracle's
 * PixelGeometry).  Out-of-grid samples are skipped rather than added as 0.
 * Build: gcc -O3 -fopenmp -ffp-contract=off.  OpenMP threads split the image rows (y); per row
 * the column values are formed first, then the voxels are visited slice by slice.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#include <omp.h>

static inline double dmin2(double a, double b) { return b < a ? b : a; }

/* Per-view state of one voxel column (x, y). */
typedef struct {
    double rs;         /* R - s */
    double W;          /* R^2 / (R - s)^2 */
    double a, b;       /* weights of the two u nodes */
    double u_edge;     /* distance of the u index from 0 and nu - 1 */
    int xl, xh;        /* the two u nodes */
    int u_ok;          /* both u nodes on the grid */
} Column;

int pxo_backproject(const double* q, int n_views, int nv, int nu, const double* cos_a,
                    const double* sin_a, double R, double scale, const double* ki, const double* p,
                    const double* xs, const double* ys, const double* zs, int nx, int ny, int nz,
                    int threads, double* vol, double* margin)
{
    if (!q || !vol || n_views < 0 || nv < 2 || nu < 2 || nx < 1 || ny < 1 || nz < 1 || threads < 1)
        return 1;
    const double ku = ki[1] - ki[0], kv = p[1] - p[0];
    for (int i = 0; i < n_views; ++i) {
        const double* qi = q + (size_t)i * nv * nu;
        const double c = cos_a[i], sn = sin_a[i];
        int failed = 0;
#pragma omp parallel num_threads(threads) reduction(| : failed)
        {
            Column* col = (Column*)malloc(sizeof(Column) * (size_t)nx);
            if (!col) failed = 1;
#pragma omp for schedule(dynamic, 1)
            for (int y = 0; y < ny; ++y) {
                if (failed) continue;
                const double Y = ys[y];
                for (int x = 0; x < nx; ++x) {
                    Column* cl = &col[x];
                    const double X = xs[x];
                    const double t = X * c + Y * sn;
                    const double s = -X * sn + Y * c;
                    const double U = R * t / (R - s);
                    cl->rs = R - s;
                    cl->W = R * R / ((R - s) * (R - s));
                    const double fu = (U - ki[0]) / ku;
                    const double xl = floor(fu), xh = ceil(fu);
                    cl->u_ok = xl >= 0.0 && xh < nu;
                    cl->a = cl->b = 0.0;
                    cl->xl = cl->xh = 0;
                    if (cl->u_ok) {
                        cl->xl = (int)xl;
                        cl->xh = (int)xh;
                        double w1 = U - ki[cl->xl];
                        const double w2 = ki[cl->xh] - U;
                        if (fabs(w1) < 1e-4 && fabs(w2) < 1e-4) w1 = 1.0;
                        cl->a = w2 / (w1 + w2);
                        cl->b = w1 / (w1 + w2);
                    }
                    cl->u_edge = dmin2(fabs(fu), fabs(fu - (nu - 1)));
                }
                for (int z = 0; z < nz; ++z) {
                    const double Rz = R * zs[z];
                    double* vrow = vol + ((size_t)z * ny + y) * nx;
                    double* mrow = margin ? margin + ((size_t)z * ny + y) * nx : NULL;
                    for (int x = 0; x < nx; ++x) {
                        const Column* cl = &col[x];
                        const double V = Rz / cl->rs;
                        const double fv = (V - p[0]) / kv;
                        if (mrow) mrow[x] = dmin2(mrow[x], dmin2(cl->u_edge, dmin2(fabs(fv), fabs(fv - (nv - 1)))));
                        if (!cl->u_ok) continue;
                        const double yl = floor(fv), yh = ceil(fv);
                        if (!(yl >= 0.0 && yh < nv)) continue;
                        double w3 = V - p[(int)yl];
                        const double w4 = p[(int)yh] - V;
                        if (fabs(w3) < 1e-4 && fabs(w4) < 1e-4) w3 = 1.0;
                        const double* r0 = qi + (size_t)yl * nu;
                        const double* r1 = qi + (size_t)yh * nu;
                        const double M = cl->a * r0[cl->xl] + cl->b * r0[cl->xh];
                        const double N = cl->a * r1[cl->xl] + cl->b * r1[cl->xh];
                        const double S = (w4 / (w3 + w4)) * M + (w3 / (w3 + w4)) * N;
                        vrow[x] += S * scale * cl->W;
                    }
                }
            }
            free(col);
        }
        if (failed) return 2;
    }
    return 0;
}
