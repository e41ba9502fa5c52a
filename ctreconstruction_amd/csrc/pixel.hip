// pixel.hip -- pixel-driven (voxel-driven, bilinear) cone-beam backprojection for gfx950.
//
// Replaces the per-view kernels CalculateInterpolGrid + Interpol2dLineargpu and the host loop
// of ConeBeam.Reconstruction (python/ConeBeamGPU.py:65-129, 303-315).  For a voxel at
// (x, y, z) and a view at angle th: t = x cos + y sin, s = -x sin + y cos,
// U = R t/(R - s), V = R z/(R - s), W = R^2/(R - s)^2 (R = source-to-axis distance), and the
// voxel gains scale * W * bilinear(Q_view; U, V) with the bilinear sample 0 outside the
// detector grid (ki, p) = ((j - (nu-1)/2), (k - (nv-1)/2)) * (du, dv) * R/SDD.
//
// Kernels:
//   px_view_kernel               cos/sin per view, FP64 -> FP32
//   px_tile_kernel               per (view, 16x16 column tile): range of the detector column
//                                coordinate and of the magnification R/(R-s) over the tile corners
//   px_fill_kernel               zeroes the per-tile "has views for the rest pass" flags
//   pixel_backproject_kernel     one workgroup per 16 x 16 x TZ voxel tile, one work-item per
//                                (x, y) column holding TZ accumulators; per view the tile's
//                                detector window is fetched a view ahead into registers, staged
//                                in LDS and every voxel reads its four samples there (two
//                                ds_read2)
//   pixel_backproject_rest_kernel   the (tile, view) pairs whose window is larger than the LDS
//                                tile, read straight from global memory and added to the volume
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "ctrecon.h"

namespace {

constexpr int kPxTile = 16;     // in-plane tile edge (voxels)
constexpr int kPxTz = 32;       // z voxels per work-item (kPxTzShort where windows are taller)
constexpr int kPxTzShort = 16;
// LDS window row pitch (floats): a window has at most 64 columns; 71 (7 mod 32) spreads the
// samples of an 8 x 8 voxel patch, which lie on a few neighbouring rows, over the banks best
// among the pitches measured (71 -> 35-36 ms, 65 -> 38 ms at 512^3 x 680 views with two windows
// per group)
constexpr int kPxPitch = 71;
constexpr int kPxCols = 64;     // window columns: one per lane of the staging loads
constexpr int kPxRows = 72;     // LDS window rows (20 KB per group)
constexpr int kPxThreads = kPxTile * kPxTile;

size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

int status_from_hip(hipError_t e) { return e == hipSuccess ? CTR_OK : CTR_ERR_HIP; }

struct PxArgs {
    int nx, ny, nz, nu, nv, n_views;
    int tiles_x, tiles_y, n_tiles_xy, accumulate;
    float x_first, y_first, z_first, dx, dy, dz;
    float R, inv_dk, u_mid, u_last, inv_dp, v_mid, v_last, scale;
};

struct PxTileArgs {
    int n_views, tiles_x, n_tiles_xy, nx, ny, tz, slabs, rows_cap;
    double x_first, y_first, z_first, dx, dy, dz, R, inv_dk, u_mid, inv_dp, v_mid;
};

__global__ void px_view_kernel(const double* __restrict__ angles, float2* __restrict__ cs,
                               int n_views) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_views) return;
    double s, c;
    sincos(angles[v], &s, &c);
    cs[v] = make_float2((float)c, (float)s);
}

// U and R/(R-s) are monotone along every line of the (x, y) plane while R - s > 0, so their
// extremes over a tile are at its corner voxel centres.  A tile that overhangs the grid is cut
// at the last voxel: the work-items past it sample at that voxel's position.
//
// The same pass decides, in FP64 and once for both backprojection kernels, whether the
// (tile, view) window fits the LDS tile for every z slab.  The main kernel sizes a window as
// floor(hi) - floor(lo) + 4 cells from float32 ends hi, lo; since floor(a) - floor(b) <=
// ceil(a - b) and float32 moves each end by far less than half a cell, that is at most
// ceil(span) + 5 for the FP64 span hi - lo.  The row span m z / dp of a slab grows with |z|, so
// the first and the last slab bound it.  A window that does not fit is marked by a negative m_lo,
// and its tile by rest[tile] = 1.
__global__ void px_tile_kernel(const double* __restrict__ angles, float4* __restrict__ tb,
                               int* __restrict__ rest, PxTileArgs a) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.n_views * a.n_tiles_xy) return;
    const int v = idx / a.n_tiles_xy;
    const int t = idx - v * a.n_tiles_xy;
    const int bx = t % a.tiles_x, by = t / a.tiles_x;
    double s, c;
    sincos(angles[v], &s, &c);
    const double xs[2] = {a.x_first + (double)(bx * kPxTile) * a.dx,
                          a.x_first + (double)min(bx * kPxTile + kPxTile - 1, a.nx - 1) * a.dx};
    const double ys[2] = {a.y_first + (double)(by * kPxTile) * a.dy,
                          a.y_first + (double)min(by * kPxTile + kPxTile - 1, a.ny - 1) * a.dy};
    double fu_lo = INFINITY, fu_hi = -INFINITY, m_lo = INFINITY, m_hi = -INFINITY;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const double tt = xs[i] * c + ys[j] * s;
            const double ss = -xs[i] * s + ys[j] * c;
            const double m = a.R / (a.R - ss);
            const double fu = tt * m * a.inv_dk + a.u_mid;
            fu_lo = fmin(fu_lo, fu);
            fu_hi = fmax(fu_hi, fu);
            m_lo = fmin(m_lo, m);
            m_hi = fmax(m_hi, m);
        }
    const double cols = ceil(fu_hi - fu_lo) + 5.0;
    double rows = 0.0;
    for (int e = 0; e < 2; ++e) {
        const double zlo = a.z_first + (double)((e ? a.slabs - 1 : 0) * a.tz) * a.dz;
        const double zhi = zlo + (double)(a.tz - 1) * a.dz;
        const double lo = (zlo < 0.0 ? m_hi : m_lo) * zlo * a.inv_dp;
        const double hi = (zhi >= 0.0 ? m_hi : m_lo) * zhi * a.inv_dp;
        rows = fmax(rows, ceil(hi - lo) + 5.0);
    }
    const bool fits = cols <= (double)kPxCols && rows <= (double)a.rows_cap;
    tb[idx] = make_float4((float)fu_lo, (float)fu_hi, fits ? (float)m_lo : -(float)m_lo, (float)m_hi);
    if (!fits) rest[t] = 1;
}

__global__ void px_fill_kernel(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The reference's inclusive row test 0 <= fv <= v_last (python/ConeBeamGPU.py:80-102) as one
// unsigned compare of the bit patterns: non-negative floats order like their bits, and negative
// values and NaN have patterns above any finite v_last >= 0.  The one value the two forms treat
// differently, -0.0, never occurs: an FMA rounds a zero sum to +0 unless both of its terms are
// -0.0, fv is an FMA onto fvb, and fvb is an FMA onto v_mid = (nv - 1) / 2 >= +0 (or an
// integer-valued float), so neither is ever -0.0.
__device__ inline bool px_row_ok(float fv, unsigned vlast_bits) { return __float_as_uint(fv) <= vlast_bits; }

// Window of a (tile, view) pair: detector columns floor(fu_lo)-1 .. floor(fu_hi)+2 and rows
// likewise from the extremes of V over the tile's z range (all values uniform in the group).
struct PxWindow {
    int c0, cols, r0, rows;
};

// V = m z is smallest at (zlo, m_lo) for zlo >= 0, else at (zlo, m_hi), and largest at
// (zhi, m_hi) for zhi >= 0, else at (zhi, m_lo); `lo_hi`/`hi_hi` say which m each end takes.
template <int ROWS>
__device__ inline PxWindow px_window(const float4 bnd, float zlo_p, float zhi_p, bool lo_hi,
                                     bool hi_hi, const PxArgs& a) {
    PxWindow w;
    w.c0 = (int)floorf(bnd.x) - 1;
    w.cols = (int)floorf(bnd.y) + 3 - w.c0;
    const float fv_lo = fmaf(lo_hi ? bnd.w : bnd.z, zlo_p, a.v_mid);
    const float fv_hi = fmaf(hi_hi ? bnd.w : bnd.z, zhi_p, a.v_mid);
    w.r0 = (int)floorf(fv_lo) - 1;
    w.rows = (int)floorf(fv_hi) + 3 - w.r0;
    // the tile pass admits only windows whose float32 size is within the LDS tile; the clamps
    // keep every LDS index inside it regardless
    w.cols = min(w.cols, kPxCols);
    w.rows = min(w.rows, ROWS);
    return w;
}

// One group per 16 x 16 x TZ voxel tile.  Per view, the window rows are fetched with one buffer
// load per (row, column) into registers a view ahead, written to the group's LDS window after a
// barrier (the previous view's reads are done) and read back by the voxels after a second one.
// One window per group rather than two lets six groups share a CU (the registers allow six,
// two 20 KB windows only four): 33.3 ms instead of 35.2 at 512^3 x 680 views.  Cells outside the
// detector get a buffer offset past the view and the range check returns 0 for them.  Views whose window does not fit (m_lo < 0 in the tile
// table) are left to pixel_backproject_rest_kernel.
template <int TZ, int ROWS>
__global__ void __launch_bounds__(kPxThreads) pixel_backproject_kernel(
    const float* __restrict__ q, float* __restrict__ vol, const float2* __restrict__ cs,
    const float4* __restrict__ tb, PxArgs a) {
    constexpr int kBuf = ROWS * kPxPitch;
    constexpr int kStage = ROWS / 4;
    __shared__ float win[kBuf];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // each wave takes an 8 x 8 quadrant of the 16 x 16 columns
    const int tx = (lane & 7) | ((wave & 1) << 3), ty = (lane >> 3) | ((wave >> 1) << 3);
    int b = blockIdx.x;
    const int bx = b % a.tiles_x;
    b /= a.tiles_x;
    const int by = b % a.tiles_y;
    const int bz = b / a.tiles_y;
    const int tile = by * a.tiles_x + bx;
    const int ix = bx * kPxTile + tx, iy = by * kPxTile + ty, iz0 = bz * TZ;
    // past the grid's x/y end a work-item samples at the last voxel (its sums are not stored);
    // past its z end the nominal positions are kept, which only lengthens the window in v
    const float x = fmaf((float)min(ix, a.nx - 1), a.dx, a.x_first);
    const float y = fmaf((float)min(iy, a.ny - 1), a.dy, a.y_first);
    const float zlo = fmaf((float)iz0, a.dz, a.z_first);
    const float zhi = fmaf((float)(iz0 + TZ - 1), a.dz, a.z_first);
    const float zlo_p = zlo * a.inv_dp, zhi_p = zhi * a.inv_dp;  // z in detector rows
    const bool lo_hi = zlo < 0.0f, hi_hi = zhi >= 0.0f;
    const size_t view_cells = (size_t)a.nu * a.nv;
    const int view_bytes = (int)(view_cells * 4);

    float acc[TZ];
    const unsigned vlast_bits = __float_as_uint(a.v_last);
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;

    float stage[kStage];
    // issue the window loads of view v into `stage` (rows wave, wave + 4, ...)
    auto fetch = [&](int v, const PxWindow& w) {
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(q + (size_t)v * view_cells), 0, view_bytes, 0x00020000);
        const int gc = w.c0 + lane;
        const unsigned lane_off =
            (lane < w.cols && (unsigned)gc < (unsigned)a.nu) ? 4u * (unsigned)gc : 0x7ffffff0u;
#pragma unroll
        for (int i = 0; i < kStage; ++i) {
            const int r = i * 4 + wave;
            if (r < w.rows) {
                const int gr = w.r0 + r;
                const unsigned row_off =
                    (unsigned)gr < (unsigned)a.nv ? 4u * (unsigned)(gr * a.nu) : 0x7ffffff0u;
                stage[i] = __uint_as_float(
                    __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)(lane_off + row_off), 0, 0));
            }
        }
    };
    auto put = [&](float* buf, const PxWindow& w) {
#pragma unroll
        for (int i = 0; i < kStage; ++i) {
            const int r = i * 4 + wave;
            if (r < w.rows) buf[r * kPxPitch + lane] = stage[i];
        }
    };

    float4 bn = tb[tile];
    PxWindow wn = px_window<ROWS>(bn, zlo_p, zhi_p, lo_hi, hi_hi, a);
    bool fits_n = bn.z > 0.0f;
    if (fits_n) fetch(0, wn);

    for (int v = 0; v < a.n_views; ++v) {
        const PxWindow w = wn;
        const bool fits = fits_n;
        float* buf = win;
        if (fits) {
            __syncthreads();
            put(buf, w);
        }
        if (v + 1 < a.n_views) {
            bn = tb[(size_t)(v + 1) * a.n_tiles_xy + tile];
            wn = px_window<ROWS>(bn, zlo_p, zhi_p, lo_hi, hi_hi, a);
            fits_n = bn.z > 0.0f;
            if (fits_n) fetch(v + 1, wn);
        }
        if (!fits) continue;
        __syncthreads();

        // this column's sample geometry (python/ConeBeamGPU.py:125-127)
        const float2 sc = cs[v];
        const float t = fmaf(x, sc.x, y * sc.y);
        const float s = fmaf(y, sc.x, -x * sc.y);
        const float m = a.R * __builtin_amdgcn_rcpf(a.R - s);  // R/(R-s), 1 ulp
        const float fu = fmaf(t * m, a.inv_dk, a.u_mid);
        const bool uok = fu >= 0.0f && fu <= a.u_last;
        const float kz = m * a.dz * a.inv_dp;
        const float fvb = fmaf(m * zlo, a.inv_dp, a.v_mid);
        // iu = floor(fu) except on the last grid column (fu = nu-1: iu = nu-2, fx = 1)
        const int iu = min((int)floorf(fu), a.nu - 2);
        const float fx = fu - (float)iu;
        // every row floor(fv(k)) must lie in [r0, r0 + rows - 2]: fv(k) is monotone in k, so
        // both ends in [r0, r0 + rows - 1) suffice (false for NaN)
        const float f_last = fmaf((float)(TZ - 1), kz, fvb);
        const float wr0 = (float)w.r0, wr1 = (float)(w.r0 + w.rows - 1);
        const bool in_win = fvb >= wr0 && fvb < wr1 && f_last >= wr0 && f_last < wr1;
        const float wc = (uok && in_win) ? m * m * a.scale : 0.0f;
        const float kz_c = in_win ? kz : 0.0f;
        const float fvb_c = in_win ? fvb : wr0;
        const int cu = min(max(iu - w.c0, 0), w.cols - 2);
        const float* base = buf + cu - w.r0 * kPxPitch;
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
            const float fv = fmaf((float)k, kz_c, fvb_c);
            const int iv = (int)floorf(fv);
            const float fy = __builtin_amdgcn_fractf(fv);
            const float* p = base + __mul24(iv, kPxPitch);
            const float q00 = p[0], q01 = p[1], q10 = p[kPxPitch], q11 = p[kPxPitch + 1];
            const float m0 = fmaf(fx, q01 - q00, q00);
            const float m1 = fmaf(fx, q11 - q10, q10);
            const float val = fmaf(fy, m1 - m0, m0);
            const float wk = px_row_ok(fv, vlast_bits) ? wc : 0.0f;
            acc[k] = fmaf(val, wk, acc[k]);
        }
    }

    if (ix < a.nx && iy < a.ny) {
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
            const int iz = iz0 + k;
            if (iz < a.nz) {
                const size_t o = ((size_t)iz * a.ny + iy) * a.nx + ix;
                vol[o] = a.accumulate ? vol[o] + acc[k] : acc[k];
            }
        }
    }
}

// The views whose window does not fit the LDS tile, for the tiles that have any: the four
// samples of a voxel come straight from global memory (indices clamped to the detector, the
// sample weighted 0 off it), and the sums are added to what pixel_backproject_kernel stored.
template <int TZ>
__global__ void __launch_bounds__(kPxThreads) pixel_backproject_rest_kernel(
    const float* __restrict__ q, float* __restrict__ vol, const float2* __restrict__ cs,
    const float4* __restrict__ tb, const int* __restrict__ rest, PxArgs a) {
    const int tid = threadIdx.x;
    const int tx = tid & (kPxTile - 1), ty = tid / kPxTile;
    int b = blockIdx.x;
    const int bx = b % a.tiles_x;
    b /= a.tiles_x;
    const int by = b % a.tiles_y;
    const int bz = b / a.tiles_y;
    const int tile = by * a.tiles_x + bx;
    if (rest[tile] == 0) return;
    const int ix = bx * kPxTile + tx, iy = by * kPxTile + ty, iz0 = bz * TZ;
    const float x = fmaf((float)min(ix, a.nx - 1), a.dx, a.x_first);
    const float y = fmaf((float)min(iy, a.ny - 1), a.dy, a.y_first);
    const float zlo = fmaf((float)iz0, a.dz, a.z_first);
    const size_t view_cells = (size_t)a.nu * a.nv;

    float acc[TZ];
    const unsigned vlast_bits = __float_as_uint(a.v_last);
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;
    for (int v = 0; v < a.n_views; ++v) {
        if (tb[(size_t)v * a.n_tiles_xy + tile].z > 0.0f) continue;
        const float2 sc = cs[v];
        const float t = fmaf(x, sc.x, y * sc.y);
        const float s = fmaf(y, sc.x, -x * sc.y);
        const float m = a.R * __builtin_amdgcn_rcpf(a.R - s);
        const float fu = fmaf(t * m, a.inv_dk, a.u_mid);
        const bool uok = fu >= 0.0f && fu <= a.u_last;
        const float kz = m * a.dz * a.inv_dp;
        const float fvb = fmaf(m * zlo, a.inv_dp, a.v_mid);
        const float* qv = q + (size_t)v * view_cells;
        const int iu = min(max((int)floorf(fu), 0), a.nu - 2);
        const float fx = __saturatef(fu - (float)iu);
        const float wc = uok ? m * m * a.scale : 0.0f;
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
            const float fv = fmaf((float)k, kz, fvb);
            const bool ok = px_row_ok(fv, vlast_bits);
            const int iv = min(max((int)floorf(fv), 0), a.nv - 2);
            const float fy = __saturatef(fv - (float)iv);
            const float* r0p = qv + (size_t)iv * a.nu + iu;
            const float q00 = r0p[0], q01 = r0p[1], q10 = r0p[a.nu], q11 = r0p[a.nu + 1];
            const float m0 = fmaf(fx, q01 - q00, q00);
            const float m1 = fmaf(fx, q11 - q10, q10);
            const float val = fmaf(fy, m1 - m0, m0);
            acc[k] = fmaf(val, ok ? wc : 0.0f, acc[k]);
        }
    }
    if (ix < a.nx && iy < a.ny) {
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
            const int iz = iz0 + k;
            if (iz < a.nz) {
                const size_t o = ((size_t)iz * a.ny + iy) * a.nx + ix;
                vol[o] += acc[k];
            }
        }
    }
}

struct PxLayout {
    size_t cs_off, tb_off, rest_off, total;
    int tiles_x, tiles_y, slabs, tz;
};

// z voxels per work-item: kPxTz when the row span of the outermost tile (largest magnification
// m_max = R/(R - r_max), its change across a tile, the largest |z|) passes the tile pass's
// test ceil(span) + 5 <= kPxRows, else kPxTzShort.  (Either way a (tile, view) window that does
// not fit is read from global memory by the rest pass.)
int px_choose_tz(const ctr_geometry* g) {
    const double xr = fabs(0.5 * (g->nx - 1) * g->dx) + fabs(g->recon_center[0]);
    const double yr = fabs(0.5 * (g->ny - 1) * g->dy) + fabs(g->recon_center[1]);
    const double m_max = g->sad / (g->sad - sqrt(xr * xr + yr * yr));
    const double dp = g->dv * g->sad / g->sdd;
    const double z_max = fabs(0.5 * (g->nz - 1) * g->dz) + fabs(g->recon_center[2]);
    const double tile_span = kPxTile * sqrt(g->dx * g->dx + g->dy * g->dy);
    const double span = (kPxTz - 1) * g->dz * m_max / dp +
                        m_max * m_max * tile_span / g->sad * z_max / dp;
    return ceil(span) + 5.0 <= kPxRows ? kPxTz : kPxTzShort;
}

PxLayout px_layout(const ctr_geometry* g) {
    PxLayout l;
    l.tiles_x = (g->nx + kPxTile - 1) / kPxTile;
    l.tiles_y = (g->ny + kPxTile - 1) / kPxTile;
    l.tz = px_choose_tz(g);
    l.slabs = (g->nz + l.tz - 1) / l.tz;
    l.cs_off = 0;
    l.tb_off = align_up((size_t)g->n_views * sizeof(float2), 256);
    l.rest_off = l.tb_off + align_up((size_t)g->n_views * l.tiles_x * l.tiles_y * sizeof(float4), 256);
    l.total = l.rest_off + align_up((size_t)l.tiles_x * l.tiles_y * sizeof(int), 256);
    return l;
}

// Checks of the pixel-driven path: the generic limits of ctr_geometry plus what the
// reference's ConeBeam geometry has (flat detector, no offset, no helix, nu, nv >= 2 so that
// ki[1] and p[1] exist) and a source outside every voxel column of the tile grid.
int check_pixel_geometry(const ctr_geometry* g) {
    if (!g) return CTR_ERR_INVALID;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->n_views <= 0) return CTR_ERR_INVALID;
    if (g->nu < 2 || g->nv < 2) return CTR_ERR_INVALID;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !(g->du > 0) || !(g->dv > 0))
        return CTR_ERR_INVALID;
    if (!(g->sad > 0) || !(g->sdd > 0)) return CTR_ERR_INVALID;
    if (g->detector_shape != CTR_DETECTOR_FLAT) return CTR_ERR_UNSUPPORTED;
    if (g->det_offset[0] != 0.0 || g->det_offset[1] != 0.0 || g->helical != 0.0)
        return CTR_ERR_UNSUPPORTED;
    if ((double)g->nx * g->ny * g->nz >= 2147483647.0) return CTR_ERR_INVALID;
    if ((double)g->n_views * g->nv * g->nu >= 2147483647.0) return CTR_ERR_INVALID;
    if (g->nx > (1 << 24) || g->ny > (1 << 24) || g->nz > (1 << 24)) return CTR_ERR_INVALID;
    const double tiles = (double)((g->nx + kPxTile - 1) / kPxTile) * ((g->ny + kPxTile - 1) / kPxTile);
    if (tiles * ((g->nz + kPxTzShort - 1) / kPxTzShort) >= 2147483647.0) return CTR_ERR_INVALID;
    if ((double)g->nu * g->nv * 4.0 >= 2147483632.0) return CTR_ERR_INVALID;  // buffer offsets
    if (tiles * g->n_views >= 2147483647.0) return CTR_ERR_INVALID;
    // farthest voxel column from the rotation axis: R - s stays >= R/1000 for every view
    const double xr = fabs(0.5 * (g->nx - 1) * g->dx) + fabs(g->recon_center[0]);
    const double yr = fabs(0.5 * (g->ny - 1) * g->dy) + fabs(g->recon_center[1]);
    if (!(g->sad - sqrt(xr * xr + yr * yr) >= 1e-3 * g->sad)) return CTR_ERR_UNSUPPORTED;
    return CTR_OK;
}

}  // namespace

size_t ctr_pixel_backproject_workspace_bytes(const ctr_geometry* g) {
    if (check_pixel_geometry(g)) return 0;
    return px_layout(g).total;
}

int ctr_pixel_backproject(const ctr_geometry* g, const double* angles, const float* q,
                          float* vol, float scale, int accumulate, void* workspace,
                          size_t workspace_bytes, void* stream) {
    int st = check_pixel_geometry(g);
    if (st) return st;
    if (!angles || !q || !vol || !workspace) return CTR_ERR_INVALID;
    if (!std::isfinite(scale)) return CTR_ERR_INVALID;
    const PxLayout l = px_layout(g);
    if (workspace_bytes < l.total) return CTR_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    float2* cs = (float2*)(ws + l.cs_off);
    float4* tb = (float4*)(ws + l.tb_off);
    int* rest = (int*)(ws + l.rest_off);
    const double mag = g->sad / g->sdd;
    const double x_first = -0.5 * (g->nx - 1) * g->dx + g->recon_center[0];
    const double y_first = -0.5 * (g->ny - 1) * g->dy + g->recon_center[1];
    const double z_first = -0.5 * (g->nz - 1) * g->dz + g->recon_center[2];
    const double inv_dk = 1.0 / (g->du * mag), inv_dp = 1.0 / (g->dv * mag);
    const int n_tiles_xy = l.tiles_x * l.tiles_y;

    hipLaunchKernelGGL(px_fill_kernel, dim3((n_tiles_xy + 255) / 256), dim3(256), 0, s, rest,
                       n_tiles_xy);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    hipLaunchKernelGGL(px_view_kernel, dim3((g->n_views + 255) / 256), dim3(256), 0, s, angles,
                       cs, g->n_views);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    PxTileArgs ta;
    ta.n_views = g->n_views;
    ta.tiles_x = l.tiles_x;
    ta.n_tiles_xy = n_tiles_xy;
    ta.nx = g->nx;
    ta.ny = g->ny;
    ta.tz = l.tz;
    ta.slabs = l.slabs;
    ta.rows_cap = kPxRows;
    ta.x_first = x_first;
    ta.y_first = y_first;
    ta.z_first = z_first;
    ta.dx = g->dx;
    ta.dy = g->dy;
    ta.dz = g->dz;
    ta.R = g->sad;
    ta.inv_dk = inv_dk;
    ta.u_mid = 0.5 * (g->nu - 1);
    ta.inv_dp = inv_dp;
    ta.v_mid = 0.5 * (g->nv - 1);
    const int n_tb = g->n_views * n_tiles_xy;
    hipLaunchKernelGGL(px_tile_kernel, dim3((n_tb + 255) / 256), dim3(256), 0, s, angles, tb,
                       rest, ta);
    if ((st = status_from_hip(hipGetLastError()))) return st;

    PxArgs a;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.nu = g->nu;
    a.nv = g->nv;
    a.n_views = g->n_views;
    a.tiles_x = l.tiles_x;
    a.tiles_y = l.tiles_y;
    a.n_tiles_xy = n_tiles_xy;
    a.accumulate = accumulate ? 1 : 0;
    a.x_first = (float)x_first;
    a.y_first = (float)y_first;
    a.z_first = (float)z_first;
    a.dx = (float)g->dx;
    a.dy = (float)g->dy;
    a.dz = (float)g->dz;
    a.R = (float)g->sad;
    a.inv_dk = (float)inv_dk;
    a.u_mid = (float)(0.5 * (g->nu - 1));
    a.u_last = (float)(g->nu - 1);
    a.inv_dp = (float)inv_dp;
    a.v_mid = (float)(0.5 * (g->nv - 1));
    a.v_last = (float)(g->nv - 1);
    a.scale = scale;
    const dim3 grid((unsigned)(n_tiles_xy * l.slabs));
    if (l.tz == kPxTz) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(pixel_backproject_kernel<kPxTz, kPxRows>), grid,
                           dim3(kPxThreads), 0, s, q, vol, cs, tb, a);
        if ((st = status_from_hip(hipGetLastError()))) return st;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(pixel_backproject_rest_kernel<kPxTz>), grid,
                           dim3(kPxThreads), 0, s, q, vol, cs, tb, rest, a);
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(pixel_backproject_kernel<kPxTzShort, kPxRows>), grid,
                           dim3(kPxThreads), 0, s, q, vol, cs, tb, a);
        if ((st = status_from_hip(hipGetLastError()))) return st;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(pixel_backproject_rest_kernel<kPxTzShort>), grid,
                           dim3(kPxThreads), 0, s, q, vol, cs, tb, rest, a);
    }
    return status_from_hip(hipGetLastError());
}
