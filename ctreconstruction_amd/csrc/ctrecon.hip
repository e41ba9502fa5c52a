// ctrecon.hip -- gfx950 kernels and C ABI for the distance-driven CT projector pair and FDK.
//
// Kernels (all FP32 arithmetic; FP64 only in the per-view geometry kernels):
//   fp_geometry_kernel          per-view DD footprint tables, FP64 -> FP32 (index space)
//   zprefix_kernel              vol[z][y][x] -> z-prefix tables [y][x][k] and [x][y][k]
//   dd_forward_kernel           one wavefront per (2 views, 4 detector columns, 63 rows)
//   dd_forward_zslab_kernel     z-slab views (non-circular geometries), FP64 loops per cell
//   bp_geometry_kernel          per-view backprojection constants, FP64 -> FP32
//   dd_backproject_tile_kernel  one workgroup per 16 x 16 x ZC voxel tile, windows in LDS
//   adj_geometry_kernel         per-view constants of the inverse central projection
//   adj_rowprefix_kernel        row prefix sums of projections times path length
//   dd_adjoint_kernel           exact transpose of the forward projector, gathered per voxel
//   spectrum_kernel             filter magnitude from the Ram-Lak cosine series, FP64
//   twiddle_kernel              Stockham twiddles, stage-major, FP64 -> FP32
//   fdk_filter_kernel           cosine weight + LDS Stockham FFT convolution, two rows per FFT
//
// Reference semantics are cited at each kernel (leehoy/CTReconstruction paths).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "ctrecon.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// ------------------------------------------------------------------------------------------
// Host-side validation
// ------------------------------------------------------------------------------------------

int check_geometry(const ctr_geometry* g) {
    if (!g) return CTR_ERR_INVALID;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->nu <= 0 || g->nv <= 0 || g->n_views <= 0)
        return CTR_ERR_INVALID;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !(g->du > 0) || !(g->dv > 0))
        return CTR_ERR_INVALID;
    if (!(g->sad > 0) || !(g->sdd > 0)) return CTR_ERR_INVALID;
    if (g->detector_shape != CTR_DETECTOR_FLAT && g->detector_shape != CTR_DETECTOR_CURVED)
        return CTR_ERR_UNSUPPORTED;
    // flat indices are 32-bit inside the kernels
    const double vox = (double)g->nx * g->ny * g->nz;
    const double cells = (double)g->n_views * g->nv * g->nu;
    if (vox >= 2147483647.0 || cells >= 2147483647.0) return CTR_ERR_INVALID;
    // grid y/z limits: z-prefix tables (ny), backprojection/adjoint (ny/4, nz/4); the
    // forward projector packs transverse cell indices into 24 bits, the backprojector
    // addresses window rows with 24-bit multiplies by nu
    if (g->n_views > 65535 || g->nv > 65535 * 4 || g->ny > 65535 || g->nz > 65535 ||
        g->nx >= (1 << 24) || g->nu >= (1 << 23))
        return CTR_ERR_INVALID;
    return CTR_OK;
}

int status_from_hip(hipError_t e) { return e == hipSuccess ? CTR_OK : CTR_ERR_HIP; }

// ------------------------------------------------------------------------------------------
// Forward projection
// ------------------------------------------------------------------------------------------

// Per-view table layout (float32, stride fp_view_stride(nu)):
//   [0,        nu+1)   alpha_e : transverse index coordinate of U-edge line e at slab 0
//   [nu+1,   2nu+2)    beta_e  : its change per slab
//   [2nu+2,  3nu+2)    T0_j    : (s_0 - S_s) / (C_s[j] - S_s), fraction source->column j at slab 0
//   [3nu+2,  4nu+2)    dT_j    : change of T per slab
//   [4nu+2,  5nu+2)    sec_j   : |d_s| * sqrt(1 + s1_j^2)
//   [5nu+2,  6nu+2)    inv_j   : 1 / (C_s[j] - S_s)
//   [6nu+2]            gamma   : (S_z - ReconCenter_z)/(-dz) + nz/2
// Row table (float32): [0, nv+1) nu_k = V2[k]/(-dz); [nv+1, 2nv+1) Vc[k] (mm).
inline int fp_view_stride(int nu) { return (int)align_up((size_t)(6 * nu + 3), 64); }

struct FpLayout {
    size_t tabs_off, rows_off, volt_off, copy_bytes, total;
};

// entries per line of the z-prefix tables: breakpoints k = 0 .. nz and copies of the total up
// to a multiple of 4, so that the pair (k, k+1) around any point of [0, nz] is inside the line
// and every line starts 16-byte aligned
int zprefix_stride(int nz) { return (nz + 2 + 3) & ~3; }

// workspace: tables, row table, then the z-prefix tables of the volume (k fastest) twice:
// [y][x][k] (gathered by y-slab views) and [x][y][k] (x-slab views)
FpLayout fp_layout(const ctr_geometry* g) {
    FpLayout L;
    // + 1024: a staged segment (up to 168 breakpoints) may read past the end of the last line
    const size_t vbytes = align_up((size_t)g->nx * g->ny * zprefix_stride(g->nz) * sizeof(float) + 1024, 256);
    L.tabs_off = 0;
    L.rows_off = align_up(L.tabs_off + (size_t)fp_view_stride(g->nu) * g->n_views * 4, 256);
    L.volt_off = align_up(L.rows_off + (size_t)(2 * g->nv + 1) * 4, 256);
    L.copy_bytes = vbytes;
    L.total = L.volt_off + 2 * vbytes;
    return L;
}

// Detector lattice of FlatDetectorConstruction / CurvedDetectorConstruction
// (python/Reconstruction_pycuda.py:113-168): in-plane position of a detector point at
// transverse coordinate u (mm, or radians on the curved arc) for a view with source S.
__device__ inline void lattice_xy(const ctr_geometry& g, double c, double s, double Sx, double Sy,
                                  double u, double* px, double* py) {
    // eu = (cos, sin, 0), ew = (sin, -cos, 0)
    if (g.detector_shape == CTR_DETECTOR_FLAT) {
        *px = Sx + u * c + g.sdd * s;
        *py = Sy + u * s - g.sdd * c;
    } else {
        double su, cu;
        sincos(u, &su, &cu);
        *px = Sx + g.sdd * su * c + g.sdd * cu * s;
        *py = Sy + g.sdd * su * s - g.sdd * cu * c;
    }
}

// One thread per (edge e in [0, nu], view).  FP64 throughout; rounded to FP32 at the store.
// Follows the host geometry of distance_forward (python/Reconstruction_pycuda.py:828-864,
// 924-935): lines through the source and the cell-edge lattice, intersection length L.
__global__ void fp_geometry_kernel(ctr_geometry g, const double* __restrict__ angles,
                                   const int32_t* __restrict__ branch, float* __restrict__ tabs,
                                   float* __restrict__ rows, int stride) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int view = blockIdx.y;
    const int nu = g.nu, nv = g.nv;
    if (view == 0 && e <= nv) {
        // row edges / centres: flat lattice  v2 = -(k - (nv-1)/2) dv + dv/2 (:160),
        //                      curved lattice v2 = -(k - (nv-1)/2) dv - dv/2 (:129)
        const double half = (g.detector_shape == CTR_DETECTOR_FLAT) ? 0.5 * g.dv : -0.5 * g.dv;
        const double v2 = -((double)e - (nv - 1) / 2.0) * g.dv + half + g.det_offset[1];
        rows[e] = (float)(v2 / (-g.dz));
        if (e < nv) rows[nv + 1 + e] = (float)(-((double)e - (nv - 1) / 2.0) * g.dv + g.det_offset[1]);
    }
    if (e > nu) return;
    const double th = angles[view];
    double s, c;
    sincos(th, &s, &c);
    const double Sx = -g.sad * s, Sy = g.sad * c;
    const double Sz = g.source_init[2] + g.helical * th / (2.0 * kPi);
    const int br = branch[view];
    float* T = tabs + (size_t)view * stride;
    if (br != CTR_BRANCH_X && br != CTR_BRANCH_Y) {  // z-slabs: dd_forward_zslab_kernel
        if (e < nu) { T[2 * nu + 2 + e] = 0.f; }
        return;
    }
    const bool xs = (br == CTR_BRANCH_X);
    // slab axis s / transverse axis t; signed steps as the reference (dy, dz negated)
    const double ds = xs ? g.dx : -g.dy;
    const double dt = xs ? -g.dy : g.dx;
    const int ns = xs ? g.nx : g.ny;
    const int nt = xs ? g.ny : g.nx;
    const double rcs = xs ? g.recon_center[0] : g.recon_center[1];
    const double rct = xs ? g.recon_center[1] : g.recon_center[0];
    const double Ss = xs ? Sx : Sy;
    const double St = xs ? Sy : Sx;
    const double s0 = rcs - (ns - 1) / 2.0 * ds;  // slab 0 centre: Plane[0] + d/2
    const double step = (g.detector_shape == CTR_DETECTOR_FLAT) ? g.du : g.du / g.sdd;
    {
        // edge e: u2 = (e - (nu-1)/2) du - du/2 + offset
        const double u2 = ((double)e - (nu - 1) / 2.0) * step - step / 2.0 + g.det_offset[0];
        double bx, by;
        lattice_xy(g, c, s, Sx, Sy, u2, &bx, &by);
        const double Bs = xs ? bx : by, Bt = xs ? by : bx;
        const double m = (Bt - St) / (Bs - Ss);
        const double t0 = St + m * (s0 - Ss);
        T[e] = (float)((t0 - rct) / dt + nt / 2.0);
        T[nu + 1 + e] = (float)(ds * m / dt);
    }
    if (e < nu) {
        const double u = ((double)e - (nu - 1) / 2.0) * step + g.det_offset[0];
        double cx, cy;
        lattice_xy(g, c, s, Sx, Sy, u, &cx, &cy);
        const double Cs = xs ? cx : cy, Ct = xs ? cy : cx;
        const double inv = 1.0 / (Cs - Ss);
        T[2 * nu + 2 + e] = (float)((s0 - Ss) * inv);
        T[3 * nu + 2 + e] = (float)(ds * inv);
        const double s1 = (St - Ct) / (Ss - Cs);
        T[4 * nu + 2 + e] = (float)(fabs(ds) * sqrt(1.0 + s1 * s1));
        T[5 * nu + 2 + e] = (float)inv;
    }
    if (e == 0) T[6 * nu + 2] = (float)((Sz - g.recon_center[2]) / (-g.dz) + g.nz / 2.0);
}

// Overlap weight (index space) of cell k with the footprint [lo, hi]; the cases
// (single cell -> 1, first, last, interior) are those of python/GPUFuncs_pycuda.py:427-452.
__device__ inline float dd_weight(int k, int k0, int k1, float lo, float hi, float inv_w,
                                  float zero_tol) {
    float w;
    if (k0 == k1) {
        w = 1.0f;
    } else {
        w = (fminf(hi, (float)(k + 1)) - fmaxf(lo, (float)k)) * inv_w;
    }
    return (w < zero_tol) ? 0.0f : w;
}

// vol[z][y][x] -> z-prefix tables t_ys[y][x][k] and t_xs[x][y][k], k = 0 .. nzp-1: entry k of
// a line is P_k = f_0 + ... + f_(k-1) (voxels f_c, 0 for c >= nz), summed in FP64 and stored
// as FP32, so that the z integral of the line from 0 to c (0 <= c <= nz) is the linear
// interpolation of P at c.  One workgroup per (y, 32 x-lines): 32 (z) x 32 (x) tiles through
// LDS, coalesced along x on read and along k on write; lanes 0..31 carry the running sums of
// the 32 lines.
__global__ void __launch_bounds__(256) zprefix_kernel(const float* __restrict__ in,
                                                      float* __restrict__ t_ys,
                                                      float* __restrict__ t_xs, int nx, int ny,
                                                      int nz, int nzp) {
    __shared__ float tile[32][33];  // [z][x]
    __shared__ float line[32][33];  // [x][z]
    const int x0 = blockIdx.x * 32, y = blockIdx.y;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    double run = 0.0;  // running sum of line x0 + threadIdx.x (threads 0..31)
    for (int z0 = 0; z0 < nzp; z0 += 32) {
        for (int r = ty; r < 32; r += 8) {
            const int z = z0 + r, x = x0 + tx;
            tile[r][tx] = (z < nz && x < nx) ? in[((size_t)z * ny + y) * nx + x] : 0.0f;
        }
        __syncthreads();
        if (threadIdx.x < 32) {
            for (int r = 0; r < 32; ++r) {
                const float f = tile[r][threadIdx.x];
                line[threadIdx.x][r] = (float)run;
                run += (double)f;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int xl = e >> 5, zl = e & 31;
            const int x = x0 + xl, z = z0 + zl;
            if (x < nx && z < nzp) {
                const float v = line[xl][zl];
                t_ys[((size_t)y * nx + x) * nzp + z] = v;
                t_xs[((size_t)x * ny + y) * nzp + z] = v;
            }
        }
        __syncthreads();
    }
}

// DD forward projection of x- and y-slab views.  Kernel semantics of distance_project_on_{y,x}3
// (python/GPUFuncs_pycuda.py:398-524): at every slab the cell's four boundary lines give a
// footprint [a1,a2] x [c1,c2] that is overlapped with the voxel grid; the reference's
// per-(slab, cell) atomicAdd becomes a register sum over slabs and one store.  Coordinates
// are in index space (plane k at k), so the same code serves increasing (x) and decreasing
// (y, z) planes.
//
// One wavefront per (two adjacent views, kFpCols adjacent detector columns, kColRows rows).
// All rays of a column share their transverse footprint at every slab (the U-edge lines depend
// on the column only), so the transverse overlaps and the lines they cover are uniform across
// the wavefront.  In a batch of kFpBatch slabs, lane 4 s + c works them out for slab n0 + s,
// column c and both views; the four lanes of slab s agree (quad swizzles) on the union of the
// eight column-views' lines and z breakpoints, and the wavefront then walks the batch slab by
// slab.  At a fine angular step adjacent views cross nearly the same lines (their rays turn by
// one step, which moves a ray's crossing with a slab by a cell or a few), so one staged segment
// serves both: the prefix data fetched from beyond L2 per view falls by about a fifth.  The two
// views of a pair are projected together when they use the same slab orientation and every
// column's lines move by at most kFpPairShift cells between them; otherwise (a change of
// orientation, a coarse step) the wavefront walks the slabs once per view.
//
// z direction: the footprint of row k at a slab is [c_k, c_k+1] with c_e = nu_e tt + gamma on
// the row edges e, and its overlap-weighted sum over a z line is the line's z integral
// I(c_k+1) - I(c_k), interpolated in the z-prefix table (the two breakpoints around the edge).
// Per slab the wavefront stages the z segment of the union of its column-views' lines in LDS
// with one to four coalesced 16-byte-per-lane loads that write LDS directly (buffer_load ...
// lds: the data never passes through VGPRs, and no ds_write is issued; lanes past the segment
// get an out-of-range offset and fetch nothing).  Lane L owns edge e0 + L of kColRows + 1 = 64
// consecutive edges and sums, in FP64 over the slabs and per column-view,
// H_e = sum_slab sum_a w_a I_a(c_e) with the transverse weights w_a pre-divided by wa * tt; row k is then (H_k+1 - H_k) / (nu_k+1 - nu_k),
// scaled by the ray's length factor.  Finished rows go through the wavefront's LDS buffer so that
// the workgroup writes each detector row's 16 columns as one 64-byte run.
// The width |c_k+1 - c_k| = |nu_k+1 - nu_k| |tt| is the reference's normaliser; a footprint
// inside one cell gets weight 1 (up to rounding) as in the reference, and cells outside the
// volume are cut by clamping c_e to [0, nz].  Transverse overlaps below zero_tol times the
// footprint width are dropped as in the reference; z overlaps are not split per cell, so the
// reference's per-cell zeroing of z weights below zero_tol (1e-6 of the footprint) has no
// counterpart.  Footprints over three cells, and unions too large for the LDS segment, take an
// explicit per-line path that reads the breakpoint pairs straight from memory.
constexpr int kFpCols = 4;              // detector columns per wavefront
constexpr int kFpViews = 2;             // adjacent views per wavefront
constexpr float kFpPairShift = 8.0f;    // most a column's lines may move between the views of a walk
constexpr int kFpBatch = 64 / kFpCols;  // slabs per batch (one lane per slab and column)
constexpr int kColRows = 63;            // detector rows per wavefront (64 lanes = its 64 row edges)
#ifndef CTR_FP_WAVES
#define CTR_FP_WAVES 4
#endif
constexpr int kFpWaves = CTR_FP_WAVES;  // wavefronts per workgroup: 16 adjacent columns, 64 bytes of a row
#ifndef CTR_FP_MIN_WAVES
#define CTR_FP_MIN_WAVES 3  // resident wavefronts per SIMD the register budget must allow
#endif
constexpr int kFpMaxLoads = 4;          // segment loads per slab
constexpr int kFpMaxQuads = 64 * kFpMaxLoads;
// View pairs per block of the workgroup order: consecutive workgroups on one XCD take kViewBlock
// adjacent view pairs of one column group, then the XCD's next column group, so the workgroups
// resident on an XCD at one time trace nearby rays and share volume lines in its L2.
#ifndef CTR_FP_VIEW_BLOCK
#define CTR_FP_VIEW_BLOCK 32
#endif
constexpr int kViewBlock = CTR_FP_VIEW_BLOCK;

// two consecutive breakpoints of a z-prefix line
struct PTaps {
    float a, b;
};
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__device__ inline float lane_f(float v, int q) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), q));
}

// value of lane ^ 1 and lane ^ 2 inside each group of four lanes (DPP quad permutes)
__device__ inline int quad_xor1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false); }
__device__ inline int quad_xor2(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false); }
__device__ inline int quad_min(int v) {
    v = min(v, quad_xor1(v));
    return min(v, quad_xor2(v));
}
__device__ inline int quad_max(int v) {
    v = max(v, quad_xor1(v));
    return max(v, quad_xor2(v));
}

// per (slab, column, view) record of a batch: packed flags (see below) and the weights
constexpr int kRecHit = 1 << 10;

__global__ void __launch_bounds__(64 * kFpWaves, CTR_FP_MIN_WAVES) dd_forward_kernel(
    const float* __restrict__ t_ys, const float* __restrict__ t_xs, float* __restrict__ proj,
    const float* __restrict__ tabs, const float* __restrict__ rows,
    const int32_t* __restrict__ branch, int nx, int ny, int nz, int nzp, int nu, int nv, int stride,
    float zero_tol, int n_views, int col_groups, int n_vblocks) {
    // workgroup b runs on XCD b % 8, which takes the column groups cg = 8 c + xcd: every XCD
    // gets an equal share of every band of detector rows (the rows outside the volume's shadow
    // are cheap).  On one XCD the local index runs over (row block, view-pair block, c, pair in
    // block), fastest last.
    const int b = blockIdx.x;
    const int xcd = b & 7, l = b >> 3;
    const int runs = (col_groups + 7) / 8;
    const int vi = l % kViewBlock;
    const int r1 = l / kViewBlock;
    const int cg = (r1 % runs) * 8 + xcd;
    const int r2 = r1 / runs;
    const int vp = (r2 % n_vblocks) * kViewBlock + vi;  // view pair
    const int rb = r2 / n_vblocks;
    const int view0 = kFpViews * vp;
    if (view0 >= n_views || cg >= col_groups) return;
    const int lane = threadIdx.x & 63;
    // wave-uniform: lets the compiler keep column data and addresses in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int j0 = (cg * kFpWaves + wave) * kFpCols;  // first column of this wavefront
    // a wavefront whose columns are all past the detector has no work but stays for the
    // workgroup's final barrier
    const bool live = j0 < nu;
    const int e0 = rb * kColRows;
    const int k = e0 + lane;  // this lane's row (lanes 0 .. kColRows-1) and edge
    const bool active = live && lane < kColRows && k < nv;
    const int kk = min(k, nv - 1);
    // the pair's views: present, slab orientation (x or y slabs; z slabs write zeros)
    int br[kFpViews];
    bool use[kFpViews];
#pragma unroll
    for (int v = 0; v < kFpViews; ++v) {
        const bool present = view0 + v < n_views;
        br[v] = present ? branch[view0 + v] : -1;
        use[v] = br[v] == CTR_BRANCH_X || br[v] == CTR_BRANCH_Y;
        if (present && !use[v] && active) {
            float* o = proj + ((size_t)(view0 + v) * nv + kk) * nu + j0;
#pragma unroll
            for (int c = 0; c < kFpCols; ++c)
                if (j0 + c < nu) o[c] = 0.0f;
        }
    }
    if (!use[0] && !use[1]) return;
    const float fnz = (float)nz;
    // z extent of the wavefront's edges
    const float nue = rows[min(k, nv)];  // this lane's edge slope
    const float nu_first = rows[e0], nu_last = rows[min(e0 + kColRows, nv)];
    // as (slab, column) worker: column j0 + cl of slab n0 + (lane >> 2)
    const int cl = lane & (kFpCols - 1);
    const int jc = min(j0 + cl, nu - 1);
    const bool col_ok = j0 + cl < nu;
    // per view: gamma (scalar, and in a vector register so that the edge fma reads one scalar
    // operand) and the lane's column's footprint lines and slab fractions
    const float* T[kFpViews];
    float gam[kFpViews], gam_v[kFpViews];
    float a_al[kFpViews], b_al[kFpViews], a_be[kFpViews], b_be[kFpViews], T0[kFpViews], dT[kFpViews];
#pragma unroll
    for (int v = 0; v < kFpViews; ++v) {
        T[v] = tabs + (size_t)min(view0 + v, n_views - 1) * stride;
        gam[v] = T[v][6 * nu + 2];
        asm volatile("v_mov_b32 %0, %1" : "=v"(gam_v[v]) : "s"(gam[v]));
        a_al[v] = T[v][jc];
        b_al[v] = T[v][jc + 1];
        a_be[v] = T[v][nu + 1 + jc];
        b_be[v] = T[v][nu + 2 + jc];
        T0[v] = T[v][2 * nu + 2 + jc];
        dT[v] = T[v][3 * nu + 2 + jc];
    }
    // walks over the slabs: both views at once when they share the slab orientation and their
    // rays cross nearly the same lines (the footprint lines of each column move by at most
    // kFpPairShift cells between the two views at the first and the last slab: a fine angular
    // step), else one view per walk (view mask bit v); a coarse step would make the union of
    // lines wide and the shared segment slower than two narrow ones
    bool together = use[0] && use[1] && br[0] == br[1];
    if (together) {
        const float fN = (float)((br[0] == CTR_BRANCH_X ? nx : ny) - 1);
        float d = fmaxf(fabsf(a_al[0] - a_al[1]), fabsf(b_al[0] - b_al[1]));
        d = fmaxf(d, fabsf(fmaf(a_be[0], fN, a_al[0]) - fmaf(a_be[1], fN, a_al[1])));
        d = fmaxf(d, fabsf(fmaf(b_be[0], fN, b_al[0]) - fmaf(b_be[1], fN, b_al[1])));
        together = !__any(col_ok && !(d <= kFpPairShift));
    }
    int walks[kFpViews], n_walks = 0;
    if (live && together) {
        walks[n_walks++] = 3;
    } else if (live) {
        if (use[0]) walks[n_walks++] = 1;
        if (use[1]) walks[n_walks++] = 2;
    }
    double H[kFpViews][kFpCols];
#pragma unroll
    for (int v = 0; v < kFpViews; ++v)
#pragma unroll
        for (int c = 0; c < kFpCols; ++c) H[v][c] = 0.0;
    // two segment buffers per wavefront: a pair of slabs is loaded straight into LDS
    // (buffer_load ... lds: no VGPRs, no ds_write) while nothing of the pair is read yet
    __shared__ uint4 segs[kFpWaves][2][kFpMaxQuads];
    __shared__ float4 recs[kFpWaves][kFpViews][64];
    const char* segs_bytes = reinterpret_cast<const char*>(&segs[0][0][0]);
    const int seg_off = wave * 2 * kFpMaxQuads * 16;  // this wavefront's segment buffers in segs
    // lanes g = 64 i + lane of segment load i: (g + 1/2) as a float, for the line index
    float gh[kFpMaxLoads];
#pragma unroll
    for (int i = 0; i < kFpMaxLoads; ++i) gh[i] = (float)(64 * i + lane) + 0.5f;
    for (int w = 0; w < n_walks; ++w) {
        const int vmask = walks[w];
        const int brw = (vmask & 1) ? br[0] : br[1];
        const bool xs = (brw == CTR_BRANCH_X);
        const float* __restrict__ src = xs ? t_xs : t_ys;
        const int ns = xs ? nx : ny;
        const int nt = xs ? ny : nx;
        const float fnt = (float)nt;
        // byte offsets inside a batch of slabs fit in 31 bits (host check)
        const int line_bytes = 4 * nzp;
        const int slab_bytes = line_bytes * nt;
        const char* __restrict__ src_bytes = reinterpret_cast<const char*>(src);
        // Can any slab hit?  The footprint lines and tt are affine in the slab index, so their
        // values at the first and last slab bound every slab's; a wavefront whose columns miss
        // the volume (transversely, or in z for all of its rows) in every view of the walk
        // only writes zeros.
        bool may_hit = false;
#pragma unroll
        for (int v = 0; v < kFpViews; ++v) {
            if (!(vmask & (1 << v))) continue;
            const float fN = (float)(ns - 1);
            const float a1N = fmaf(a_be[v], fN, a_al[v]), a2N = fmaf(b_be[v], fN, b_al[v]);
            const float lo = fminf(fminf(a_al[v], b_al[v]), fminf(a1N, a2N));
            const float hi = fmaxf(fmaxf(a_al[v], b_al[v]), fmaxf(a1N, a2N));
            const float ttN = fmaf(dT[v], fN, T0[v]);
            const float z1 = fmaf(nu_first, T0[v], gam[v]), z2 = fmaf(nu_last, T0[v], gam[v]);
            const float z3 = fmaf(nu_first, ttN, gam[v]), z4 = fmaf(nu_last, ttN, gam[v]);
            const float zmin = fminf(fminf(z1, z2), fminf(z3, z4)), zmax = fmaxf(fmaxf(z1, z2), fmaxf(z3, z4));
            may_hit |= col_ok && hi >= -1.0f && lo <= fnt + 1.0f && zmax >= -1.0f && zmin <= fnz + 1.0f;
        }
        const int n_end = __ballot(may_hit) ? ns : 0;
        for (int n0 = 0; n0 < n_end; n0 += kFpBatch) {
            // the batch's slabs through one buffer resource: a line is addressed by a scalar byte
            // offset, the lane's quad by a vector one
            const char* batch = src_bytes + (size_t)n0 * slab_bytes;
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                (void*)batch, 0, (int)((size_t)min(kFpBatch, ns - n0) * slab_bytes), 0x00020000);
            // lane 4 s + c: slab n0 + s, column j0 + c.  Packed record p[v] of view v: first line
            // relative to the slab's union (bits 0-7), hit (bit 10).  Union u1: first line (bits
            // 0-23), lines staged (24-31); u2: first z breakpoint (0-15), quads per line (16-23),
            // walk (24-27: 0 no column-view hits, 1-4 staged with that many segment loads and two
            // lines per column, 5-8 the same with three lines per column, 9 explicit).
            int p[kFpViews];
            int u1 = 0, u2 = 0;
            {
                const int n = n0 + (lane >> 2);
                const float fn = (float)n;
                int kas_u = 0x7fffffff, kend_u = -1, zlo_u = 0x7fffffff, zhi_u = -1, da_u = -1;
                int kas_v[kFpViews], da_v[kFpViews];
#pragma unroll
                for (int v = 0; v < kFpViews; ++v) {
                    float tt = 1.0f, w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
                    int kas = 0x7fffffff, da = -1;
                    if ((vmask & (1 << v)) && col_ok && n < ns) {
                        const float a1 = fmaf(a_be[v], fn, a_al[v]), a2 = fmaf(b_be[v], fn, b_al[v]);
                        const float alo = fminf(a1, a2), ahi = fmaxf(a1, a2);
                        const int ks = max((int)floorf(alo), 0), ke = min((int)floorf(ahi), nt - 1);
                        tt = fmaf(dT[v], fn, T0[v]);
                        const float za = fmaf(nu_first, tt, gam[v]), zb = fmaf(nu_last, tt, gam[v]);
                        const bool z_hit = fmaxf(za, zb) >= -0.01f && fminf(za, zb) <= fnz + 0.01f;
                        if (ks <= ke && z_hit) {
                            da = ke - ks;
                            const float wa = ahi - alo;
                            // overlaps with cells ks .. ks+2 (floor(max(alo, 0)) == ks); negative
                            // values past the footprint fail the threshold test
                            const float lo = fmaxf(alo, 0.0f), hi = fminf(ahi, fnt);
                            const float s1 = (float)ks + 1.0f;
                            float o0 = fminf(hi, s1) - lo, o1 = fminf(hi, s1 + 1.0f) - s1, o2 = hi - (s1 + 1.0f);
                            const float tl = zero_tol * wa;
                            o0 = (o0 >= tl) ? o0 : 0.0f;
                            o1 = (o1 >= tl) ? o1 : 0.0f;
                            o2 = (o2 >= tl) ? o2 : 0.0f;
                            const float r = __builtin_amdgcn_rcpf(wa * tt);
                            w0 = o0 * r;
                            w1 = o1 * r;
                            w2 = o2 * r;
                            // breakpoints of the wavefront's edges: floor of the clamped end
                            // edges, the same arithmetic as each lane's (the edges are monotone
                            // in the lane)
                            const int kf = (int)floorf(__builtin_amdgcn_fmed3f(za, 0.0f, fnz));
                            const int kl = (int)floorf(__builtin_amdgcn_fmed3f(zb, 0.0f, fnz));
                            kas = ks;
                            kas_u = min(kas_u, ks);
                            // every column reads two (three) consecutive lines from its first
                            // one; lines past its footprint have weight 0 but are staged, so
                            // that every read is of staged data
                            kend_u = max(kend_u, ks + 2);
                            zlo_u = min(zlo_u, min(kf, kl));
                            zhi_u = max(zhi_u, max(kf, kl));
                            da_u = max(da_u, da);
                        }
                    }
                    recs[wave][v][lane] = make_float4(tt, w0, w1, w2);
                    kas_v[v] = kas;
                    da_v[v] = da;
                }
                const int maxda = quad_max(da_u);
                const int L0 = quad_min(kas_u);
                // two lines per column unless a footprint covers three cells
                const int L1 = quad_max(kend_u) - (maxda >= 2 ? 0 : 1);
                const int zl = quad_min(zlo_u), zh = quad_max(zhi_u);
#pragma unroll
                for (int v = 0; v < kFpViews; ++v) p[v] = 0;
                if (maxda >= 0) {
                    // the segment starts at a multiple of 4 breakpoints (16-byte aligned loads)
                    // and holds breakpoints kseg .. zh + 1; past the end of a line it reads the
                    // next one (or the table's slack, or nothing past the batch), which no weight
                    // uses
                    const int kseg = zl & ~3;
                    const int nq = (zh - kseg + 5) >> 2;
                    const int nlines = L1 - L0 + 1;
                    const int nl = (nlines * nq + 63) >> 6;
                    const bool staged = maxda <= 2 && nl <= kFpMaxLoads;
                    const int walk = !staged ? 9 : (maxda >= 2 ? 4 + nl : nl);
                    u1 = staged ? (L0 | (nlines << 24)) : 0;
                    u2 = kseg | (min(nq, 255) << 16) | (walk << 24);  // kseg <= nz < 2^16
                    // staged: kas - L0 < nlines <= kFpMaxQuads fits its 8 bits
#pragma unroll
                    for (int v = 0; v < kFpViews; ++v)
                        if (da_v[v] >= 0) p[v] = (staged ? kas_v[v] - L0 : 0) | kRecHit;
                }
            }
            const uint64_t heads = __ballot(cl == 0);
            // staged slabs, software-pipelined in pairs: both slabs' segments are loaded into the
            // wavefront's two LDS buffers, then both are summed; the next pair's loads go out only
            // after every read of this pair has returned (its sums use them)
            auto walk_slabs = [&](auto nl_c, auto lines_c, uint64_t mask) {
                constexpr int NL = decltype(nl_c)::value;        // segment loads per slab
                constexpr int LINES = decltype(lines_c)::value;  // lines read per column
                struct Set {
                    int q, b;  // slab, LDS buffer
                };
                auto issue = [&](int q, int b, Set& S) {
                    S.q = q;
                    S.b = b;
                    const int U1 = __builtin_amdgcn_readlane(u1, 4 * q), U2 = __builtin_amdgcn_readlane(u2, 4 * q);
                    const int L0 = U1 & 0xffffff, nlines = (unsigned)U1 >> 24;
                    const int kseg = U2 & 0xffff, nq = (U2 >> 16) & 0xff;
                    // quad g of the segment: line g / nq (the quotient from an FP32 reciprocal:
                    // (g + 1/2)/nq is at least 1/(2 nq) from an integer and its rounding error
                    // below 1e-4, so the floor is exact for g < 256, nq < 256), quad g - line nq,
                    // at byte line * line_bytes + 16 quad = 16 g + line (line_bytes - 16 nq)
                    const float inv_nq = __builtin_amdgcn_rcpf((float)nq);
                    const int step = line_bytes - 16 * nq;
                    const int soff = q * slab_bytes + L0 * line_bytes + 4 * kseg;
#pragma unroll
                    for (int i = 0; i < NL; ++i) {
                        const int ln = (int)(gh[i] * inv_nq);
                        // quads past the segment get an offset beyond the buffer: the range
                        // check returns 0 for them without a memory access
                        const int voff = ln < nlines ? 16 * (64 * i + lane) + __mul24(ln, step) : 0x7ffffff0;
                        // lane L's 16 bytes land at quad 64 i + L of the buffer
                        __builtin_amdgcn_raw_ptr_buffer_load_lds(
                            rsrc, (__attribute__((address_space(3))) void*)&segs[wave][b][64 * i], 16, voff, soff, 0, 0);
                    }
                };
                auto consume = [&](const Set& S) {
                    const int buf_off = seg_off + S.b * kFpMaxQuads * 16;
                    const int U2 = __builtin_amdgcn_readlane(u2, 4 * S.q);
                    const int kseg = U2 & 0xffff, nq16 = (U2 >> 12) & 0xff0;
#pragma unroll
                    for (int v = 0; v < kFpViews; ++v) {
                        if (!(vmask & (1 << v))) continue;
                        const float4* rec = recs[wave][v];
                        float4 r[kFpCols];
#pragma unroll
                        for (int c = 0; c < kFpCols; ++c) r[c] = rec[4 * S.q + c];
                        unsigned lb[kFpCols];
                        float fr[kFpCols];
#pragma unroll
                        for (int c = 0; c < kFpCols; ++c) {
                            const int pc = __builtin_amdgcn_readlane(p[v], 4 * S.q + c);
                            // byte of the column's first line in the segment buffers, less the
                            // segment's first breakpoint (scalar): one shift-add per lane
                            const unsigned sbase = (unsigned)(buf_off + (pc & 0xff) * nq16 - 4 * kseg);
                            const float cz = __builtin_amdgcn_fmed3f(fmaf(nue, r[c].x, gam_v[v]), 0.0f, fnz);
                            const float fl = floorf(cz);
                            fr[c] = cz - fl;
                            // a column that misses this slab reads the segment's start with
                            // weight 0
                            lb[c] = (pc & kRecHit) ? 4u * (unsigned)fl + sbase : (unsigned)buf_off;
                        }
                        PTaps t[kFpCols][LINES];
#pragma unroll
                        for (int c = 0; c < kFpCols; ++c)
#pragma unroll
                            for (int a = 0; a < LINES; ++a)
                                t[c][a] = *reinterpret_cast<const PTaps*>(segs_bytes + lb[c] + a * nq16);
#pragma unroll
                        for (int c = 0; c < kFpCols; ++c) {
                            const float wgt[3] = {r[c].y, r[c].z, r[c].w};
                            float sum = wgt[0] * fmaf(fr[c], t[c][0].b - t[c][0].a, t[c][0].a);
#pragma unroll
                            for (int a = 1; a < LINES; ++a)
                                sum = fmaf(wgt[a], fmaf(fr[c], t[c][a].b - t[c][a].a, t[c][a].a), sum);
                            H[v][c] += (double)sum;
                        }
                    }
                };
                // the rare walks (three or four segment loads, or three lines per column) go one
                // slab at a time (pairing them measured slower)
                constexpr bool kPair = NL <= 2 && LINES == 2;
                while (mask) {
                    Set S0, S1;
                    const int q0 = __builtin_ctzll(mask) >> 2;
                    mask &= mask - 1;
                    issue(q0, 0, S0);
                    if (kPair && mask) {
                        const int q1 = __builtin_ctzll(mask) >> 2;
                        mask &= mask - 1;
                        issue(q1, 1, S1);
                        __builtin_amdgcn_sched_barrier(0);
                        consume(S0);
                        consume(S1);
                    } else {
                        consume(S0);
                    }
                }
            };
            const int wk = (u2 >> 24) & 15;
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            using I3 = std::integral_constant<int, 3>;
            using I4 = std::integral_constant<int, 4>;
            uint64_t mk;
            if ((mk = __ballot(wk == 1) & heads)) walk_slabs(I1{}, I2{}, mk);
            if ((mk = __ballot(wk == 2) & heads)) walk_slabs(I2{}, I2{}, mk);
            if ((mk = __ballot(wk == 3) & heads)) walk_slabs(I3{}, I2{}, mk);
            if ((mk = __ballot(wk == 4) & heads)) walk_slabs(I4{}, I2{}, mk);
            if ((mk = __ballot(wk == 5) & heads)) walk_slabs(I1{}, I3{}, mk);
            if ((mk = __ballot(wk == 6) & heads)) walk_slabs(I2{}, I3{}, mk);
            if ((mk = __ballot(wk == 7) & heads)) walk_slabs(I3{}, I3{}, mk);
            if ((mk = __ballot(wk == 8) & heads)) walk_slabs(I4{}, I3{}, mk);
            // explicit slabs: every transverse cell of each column-view's footprint, pairs read
            // from memory
            uint64_t m2 = __ballot(wk == 9) & heads;
            while (m2) {
                const int q = __builtin_ctzll(m2) >> 2;
                m2 &= m2 - 1;
                const float fn = (float)(n0 + q);
#pragma unroll
                for (int v = 0; v < kFpViews; ++v) {
                    if (!(vmask & (1 << v))) continue;
#pragma unroll
                    for (int c = 0; c < kFpCols; ++c) {
                        const int pc = __builtin_amdgcn_readlane(p[v], 4 * q + c);
                        if (!(pc & kRecHit)) continue;
                        const float4 r = recs[wave][v][4 * q + c];
                        const float cz = __builtin_amdgcn_fmed3f(fmaf(nue, r.x, gam_v[v]), 0.0f, fnz);
                        const float fl = floorf(cz);
                        const float fr = cz - fl;
                        const int off = 4 * (int)fl;
                        const float a1 = fmaf(lane_f(a_be[v], c), fn, lane_f(a_al[v], c));
                        const float a2 = fmaf(lane_f(b_be[v], c), fn, lane_f(b_al[v], c));
                        const float alo = fminf(a1, a2), ahi = fmaxf(a1, a2);
                        const float wa = ahi - alo;
                        const int kas = max((int)floorf(alo), 0), kae = min((int)floorf(ahi), nt - 1);
                        float sum = 0.0f;
                        for (int ka = kas; ka <= kae; ++ka) {
                            const float ov = fminf(ahi, (float)(ka + 1)) - fmaxf(alo, (float)ka);
                            if (ov < zero_tol * wa) continue;
                            const v2u pv = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, q * slab_bytes + ka * line_bytes, 0);
                            const float pa = __uint_as_float(pv.x), pb = __uint_as_float(pv.y);
                            sum = fmaf(ov, fmaf(fr, pb - pa, pa), sum);
                        }
                        H[v][c] += (double)(sum * __builtin_amdgcn_rcpf(wa * r.x));
                    }
                }
            }
        }
    }
    // row k = edges k and k+1: the next lane's sum, times the ray's length factor.  The rows
    // leave through LDS: each wavefront parks its four columns per row in its own segment buffer
    // ([view][row][column]), and after one barrier the workgroup writes each row's 16 columns as
    // one 64-byte run, so that no cache line is written back in pieces by wavefronts that finish
    // at different times.
    const float vc = rows[nv + 1 + kk];
    const float dnu = rows[kk + 1] - rows[kk];
    float4* __restrict__ stage = reinterpret_cast<float4*>(&segs[wave][0][0]);
#pragma unroll
    for (int v = 0; v < kFpViews; ++v) {
        if (!use[v]) continue;
        float res[kFpCols];
#pragma unroll
        for (int c = 0; c < kFpCols; ++c) {
            const double Hn = __shfl_down(H[v][c], 1);
            const int jj = min(j0 + c, nu - 1);
            const float sec = T[v][4 * nu + 2 + jj], inv = T[v][5 * nu + 2 + jj];
            const float s2 = vc * inv;
            res[c] = (float)(Hn - H[v][c]) * (sec * sqrtf(1.0f + s2 * s2) / dnu);
        }
        stage[v * 64 + lane] = make_float4(res[0], res[1], res[2], res[3]);
    }
    __syncthreads();
    // item (v, r, q): view v, row e0 + r, the 4 columns of wavefront q; consecutive threads take
    // the wavefronts of one row
    const int jw = cg * kFpWaves * kFpCols;  // the workgroup's first column
    const bool vec_ok = (nu & 3) == 0 && (reinterpret_cast<uintptr_t>(proj) & 15) == 0;
    for (int it = threadIdx.x; it < kFpViews * kColRows * kFpWaves; it += 64 * kFpWaves) {
        const int q = it % kFpWaves;
        const int r = (it / kFpWaves) % kColRows;
        const int v = it / (kFpWaves * kColRows);
        const int row = e0 + r, jq = jw + q * kFpCols;
        if (!use[v] || row >= nv || jq >= nu) continue;
        const float4 val = reinterpret_cast<const float4*>(&segs[q][0][0])[v * 64 + r];
        float* __restrict__ o = proj + ((size_t)(view0 + v) * nv + row) * nu + jq;
        if (vec_ok && jq + kFpCols <= nu) {
            *reinterpret_cast<float4*>(o) = val;
        } else {
            const float vv[kFpCols] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int c = 0; c < kFpCols; ++c)
                if (jq + c < nu) o[c] = vv[c];
        }
    }
}

// z-slab views (|S_z - D_z| dominates the source-detector offset; no circular or helical orbit
// reaches it): kernel semantics of distance_project_on_z3 (python/GPUFuncs_pycuda.py:526-587)
// with the host tables of distance_forward (python/Reconstruction_pycuda.py:828-864, 993-1003).
// The slabs are z planes; at each slab the cell's four boundary lines give a footprint
// [x1, x2] x [y1, y2], and every (x, y) cell it covers adds voxel * w_x * w_y, w the reference's
// overlap weights (x planes increasing, y planes decreasing; weights below zero_tol dropped),
// the sum times the cell's intersection length L.  A rare branch, so plain FP64 loops: a few
// workgroups per view stride over its cells and return at once for x- and y-slab views.
constexpr int kZslabGroups = 16;  // workgroups per view

__global__ void __launch_bounds__(256) dd_forward_zslab_kernel(ctr_geometry g, const double* __restrict__ angles,
                                                               const int32_t* __restrict__ branch,
                                                               const float* __restrict__ vol,
                                                               float* __restrict__ proj, double zero_tol) {
    const int view = blockIdx.y;
    if (branch[view] != CTR_BRANCH_Z) return;
    const int nu = g.nu, nv = g.nv, nx = g.nx, ny = g.ny, nz = g.nz;
    const double th = angles[view];
    double s, c;
    sincos(th, &s, &c);
    const double Sx = -g.sad * s, Sy = g.sad * c;
    const double Sz = g.source_init[2] + g.helical * th / (2.0 * kPi);
    const bool flat = g.detector_shape == CTR_DETECTOR_FLAT;
    const double step = flat ? g.du : g.du / g.sdd;
    const double half = flat ? 0.5 * g.dv : -0.5 * g.dv;  // row edges: +dv/2 flat, -dv/2 curved
    // edge planes (python/Reconstruction_pycuda.py:792-797): x increasing, y and z decreasing
    const double ddx = g.dx, ddy = -g.dy, ddz = -g.dz;
    auto xplane = [&](int i) { return g.recon_center[0] + ((double)i - (nx - 1) / 2.0) * ddx - ddx / 2.0; };
    auto yplane = [&](int i) { return g.recon_center[1] + ((double)i - (ny - 1) / 2.0) * ddy - ddy / 2.0; };
    auto zplane = [&](int i) { return g.recon_center[2] + ((double)i - (nz - 1) / 2.0) * ddz - ddz / 2.0; };
    const double x0 = xplane(0), y0 = yplane(0);
    // one DD weight (python/GPUFuncs_pycuda.py:553-575): k of [ks, ke] against the footprint
    // [lo, hi] of width w; `dec`: decreasing planes p(k) > p(k+1)
    auto weight = [&](int kk, int ks, int ke, double lo, double hi, double w, double pk, double pk1, double d,
                      bool dec) {
        double r;
        if (ks == ke) r = 1.0;
        else if (kk == ks) r = dec ? (hi - pk1) / w : (pk1 - lo) / w;
        else if (kk == ke) r = dec ? (pk - lo) / w : (hi - pk) / w;
        else r = fabs(d) / w;
        return fabs(r) < zero_tol ? 0.0 : r;
    };
    for (int cell = blockIdx.x * 256 + threadIdx.x; cell < nu * nv; cell += kZslabGroups * 256) {
        const int k = cell / nu, j = cell - k * nu;
        // column edges j, j+1 and centre (in-plane), row edges k, k+1 and centre (z)
        const double ue1 = ((double)j - (nu - 1) / 2.0) * step - step / 2.0 + g.det_offset[0];
        const double ue2 = ((double)(j + 1) - (nu - 1) / 2.0) * step - step / 2.0 + g.det_offset[0];
        const double uc = ((double)j - (nu - 1) / 2.0) * step + g.det_offset[0];
        double bx1, by1, bx2, by2, cx, cy;
        lattice_xy(g, c, s, Sx, Sy, ue1, &bx1, &by1);
        lattice_xy(g, c, s, Sx, Sy, ue2, &bx2, &by2);
        lattice_xy(g, c, s, Sx, Sy, uc, &cx, &cy);
        const double cz = Sz + (-((double)k - (nv - 1) / 2.0) * g.dv + g.det_offset[1]);
        const double bz1 = Sz + (-((double)k - (nv - 1) / 2.0) * g.dv + half + g.det_offset[1]);
        const double bz2 = Sz + (-((double)(k + 1) - (nv - 1) / 2.0) * g.dv + half + g.det_offset[1]);
        // footprint lines t = a z + b through the source: x from the column edges at the row
        // centre, y from the column centre at the row edges
        const double aU1 = (Sx - bx1) / (Sz - cz), bU1 = -aU1 * Sz + Sx;
        const double aU2 = (Sx - bx2) / (Sz - cz), bU2 = -aU2 * Sz + Sx;
        const double aV1 = (Sy - cy) / (Sz - bz2), bV1 = -aV1 * Sz + Sy;
        const double aV2 = (Sy - cy) / (Sz - bz1), bV2 = -aV2 * Sz + Sy;
        const double s1 = (Sx - cx) / (Sz - cz), s2 = (Sy - cy) / (Sz - cz);
        const double L = fabs(ddz) * sqrt(1.0 + s1 * s1) * sqrt(1.0 + s2 * s2);
        double acc = 0.0;
        for (int n = 0; n < nz; ++n) {
            const double zc = zplane(n) + ddz / 2.0;
            const double t1 = aU1 * zc + bU1, t2 = aU2 * zc + bU2;
            const double v1 = aV1 * zc + bV1, v2 = aV2 * zc + bV2;
            const int ix1 = (int)floor((t1 - x0) / ddx), ix2 = (int)floor((t2 - x0) / ddx);
            const int iy1 = (int)floor((v1 - y0) / ddy), iy2 = (int)floor((v2 - y0) / ddy);
            const int sx = min(ix1, ix2), ex = max(ix1, ix2), sy = min(iy1, iy2), ey = max(iy1, iy2);
            const double tlo = fmin(t1, t2), thi = fmax(t1, t2), wt = fabs(t1 - t2);
            const double vlo = fmin(v1, v2), vhi = fmax(v1, v2), wv = fabs(v1 - v2);
            const float* slab = vol + (size_t)n * nx * ny;
            for (int kx = max(sx, 0); kx <= min(ex, nx - 1); ++kx) {
                const double w1 = weight(kx, sx, ex, tlo, thi, wt, xplane(kx), xplane(kx + 1), ddx, false);
                for (int ly = max(sy, 0); ly <= min(ey, ny - 1); ++ly) {
                    const double w2 = weight(ly, sy, ey, vlo, vhi, wv, yplane(ly), yplane(ly + 1), ddy, true);
                    acc += (double)slab[(size_t)ly * nx + kx] * w1 * w2;
                }
            }
        }
        proj[((size_t)view * nv + k) * nu + j] = (float)(acc * L);
    }
}

// ------------------------------------------------------------------------------------------
// Backprojection
// ------------------------------------------------------------------------------------------

constexpr int kBpViewStride = 16;  // floats per view in the backprojection table

// Per-view constants of flat_distance_backproj_arb2's param vector
// (python/Reconstruction_pycuda.py:405-407, 421-423), in FP64, stored as FP32:
//   [0] cos th  [1] sin th  [2] Sx  [3] Sy  [4] Sz
//   [5] Au [6] Bu : u index = Au - Bu * slope_u          (flat)
//                   u index = Au - Bu * atan(slope_u)    (curved, numba units)
//   [7] Av [8] Bv : v index = Av - Bv * slope_v
//   [9] SAD
__global__ void bp_geometry_kernel(ctr_geometry g, const double* __restrict__ angles,
                                   float* __restrict__ vt, int mode) {
    const int view = blockIdx.x * blockDim.x + threadIdx.x;
    if (view >= g.n_views) return;
    const double th = angles[view];
    double s, c;
    sincos(th, &s, &c);
    const double Sx = g.source_init[0], Sy = g.source_init[1];
    const double Sz = g.source_init[2] + g.helical * th / (2.0 * kPi);
    const double Dy = g.detector_init[1];
    const double du = (g.detector_shape == CTR_DETECTOR_FLAT) ? g.du : g.du / g.sdd;
    const double dv = -g.dv;
    // reference grid (:353-354) or true edges (:356-357)
    const double cu = (mode == CTR_BP_FDK_EDGE) ? g.nu / 2.0 : (g.nu - 1) / 2.0;
    const double cv = (mode == CTR_BP_FDK_EDGE) ? g.nv / 2.0 : (g.nv - 1) / 2.0;
    const double ki0 = -cu * du + g.det_offset[0];
    const double p0 = -cv * dv + g.det_offset[1];
    // coord = slope * Dy + (S - slope * Sy) = S - slope * (Sy - Dy)
    const double span = Sy - Dy;
    float* T = vt + (size_t)view * kBpViewStride;
    T[0] = (float)c;
    T[1] = (float)s;
    T[2] = (float)Sx;
    T[3] = (float)Sy;
    T[4] = (float)Sz;
    if (g.detector_shape == CTR_DETECTOR_FLAT) {
        T[5] = (float)((Sx - ki0) / du);
        T[6] = (float)(span / du);
    } else {
        T[5] = (float)(-ki0 / du);
        T[6] = (float)(1.0 / du);
    }
    T[7] = (float)((Sz - p0) / dv);
    T[8] = (float)(span / dv);
    T[9] = (float)g.sad;
}

// FDK-weighted DD backprojection.  Kernel semantics of flat_distance_backproj_arb2
// (python/GPUFuncs_pycuda.py:837-956) and curved_distance_backproj_arb
// (python/GPUFuncs_numba.py:537-685): the 8 voxel corners are rotated by the view angle, the
// min/max slopes through the unrotated source give the u- and v-footprints on the detector
// plane, which are overlapped with the detector grid and weighted by SAD^2/(SAD - y_c)^2.
// With angle2 = 0 the u-footprint depends only on the voxel column, so it is computed once
// per (column, view) and reused along z; the reference's per-tap atomicAdd becomes a
// register accumulation over all views and one store.
//
// Organisation: a workgroup owns a 16 x 16 x ZC voxel tile (ZC = 32); for every view it stages
// the detector window that the tile's bounding box projects to in LDS (double-buffered, one
// barrier per view).  Each work-item computes the u-footprint, u-overlap weights and FDK weight
// of its voxel column once per view for all ZC voxels, which form two sub-chunks of 16.  Per
// sub-chunk it filters every window row it needs in u once and keeps the running sum of those
// rows (a row integral C(v), piecewise linear in the index-space coordinate v) in its own LDS
// column.  The v-overlap-weighted sum of rows for a voxel whose footprint is [v_lo, v_hi] is then
// (C(v_hi) - C(v_lo)) / (v_hi - v_lo): the reference's weights are exactly the overlap lengths
// of the footprint with the rows, divided by its width.  Weight zeroing (numba variant) and
// sub-chunks spanning more than kMaxRows rows take the explicit per-row path instead.  Tiles
// are assigned so that each XCD holds contiguous z-chunks, whose windows share detector rows in
// that XCD's L2.  A window larger than the LDS buffer is gathered from global memory.
constexpr int kTileX = 16, kTileY = 16;
#ifndef CTR_BP_WIN_CAP
#define CTR_BP_WIN_CAP 1024
#endif
constexpr int kWinCap = CTR_BP_WIN_CAP;  // floats per LDS window buffer
#ifndef CTR_BP_WIN_RING
#define CTR_BP_WIN_RING 64
#endif
constexpr int kWinRing = CTR_BP_WIN_RING;  // views in the LDS ring of detector windows (power of 2)
constexpr int kMaxUTaps = 3;           // u taps held in registers (more are looped)
constexpr int kPrefetch = (kWinCap + 255) / 256;  // window cells per thread
#ifndef CTR_BP_ZC
#define CTR_BP_ZC 32
#endif
#ifndef CTR_BP_MAX_ROWS
#define CTR_BP_MAX_ROWS 19
#endif
constexpr int kBpZC = CTR_BP_ZC;            // voxels along z per work-item (and tile)
constexpr int kBpSub = 16;                  // voxels per sub-chunk: the unit of one row integral
static_assert(kBpZC % kBpSub == 0, "a work-item's voxels split into whole sub-chunks");
constexpr int kMaxRows = CTR_BP_MAX_ROWS;   // rows per sub-chunk on the row-integral path

struct BpView {
    float c, s, Sx, Sy, Sz, Au, Bu, Av, Bv, R;
};

__device__ inline BpView load_view(const float* __restrict__ vt, int view) {
    const float* T = vt + (size_t)view * kBpViewStride;
    BpView v;
    v.c = T[0]; v.s = T[1]; v.Sx = T[2]; v.Sy = T[3]; v.Sz = T[4];
    v.Au = T[5]; v.Bu = T[6]; v.Av = T[7]; v.Bv = T[8]; v.R = T[9];
    return v;
}

// u index coordinate of the ray through the source with in-plane slope `su`.  The footprint
// arithmetic below is written with explicit fmaf (nothing left for the compiler to contract):
// the fast and the completing pass decide from the same values which (work-item, view) pairs
// are the common case, so both must round them identically.
__device__ inline float u_index(const BpView& v, float su, int curved) {
    return fmaf(-v.Bu, curved ? atanf(su) : su, v.Au);
}

// Detector window (first row, rows, first column, columns) covering the projection of the box
// [xa, xb] x [ya, yb] x [za, zb]: the footprint of a box is the hull of its corners' footprints;
// one cell of margin absorbs rounding differences against the per-voxel computation.
__device__ inline void tile_window(const BpView& v, float xa, float xb, float ya, float yb,
                                   float za, float zb, int curved, int nu, int nv, int& wv0,
                                   int& wvn, int& wu0, int& wun) {
    float smin = 3.0e38f, smax = -3.0e38f, rmin = 3.0e38f, rmax = -3.0e38f;
#pragma unroll
    for (int corner = 0; corner < 4; ++corner) {
        const float x = (corner & 1) ? xb : xa;
        const float y = (corner & 2) ? yb : ya;
        const float xr = fmaf(x, v.c, y * v.s);
        const float yr = fmaf(y, v.c, -(x * v.s));
        const float r = __builtin_amdgcn_rcpf(v.Sy - yr);
        const float su = (v.Sx - xr) * r;
        smin = fminf(smin, su);
        smax = fmaxf(smax, su);
        rmin = fminf(rmin, r);
        rmax = fmaxf(rmax, r);
    }
    const float f1 = u_index(v, smin, curved), f2 = u_index(v, smax, curved);
    const int u0 = max((int)floorf(fminf(f1, f2)) - 1, 0);
    const int u1 = min((int)floorf(fmaxf(f1, f2)) + 1, nu - 1);
    const float ehi = v.Sz - za, elo = v.Sz - zb;  // za <= zb
    const float smx = ehi * (ehi >= 0.0f ? rmax : rmin);
    const float smn = elo * (elo >= 0.0f ? rmin : rmax);
    const float g1 = fmaf(-v.Bv, smn, v.Av), g2 = fmaf(-v.Bv, smx, v.Av);
    const int v0 = max((int)floorf(fminf(g1, g2)) - 1, 0);
    const int v1 = min((int)floorf(fmaxf(g1, g2)) + 1, nv - 1);
    wu0 = u0;
    wun = max(u1 - u0 + 1, 0);  // 0: the box projects outside the detector
    wv0 = v0;
    wvn = max(v1 - v0 + 1, 0);
}

// 0: window empty (skip the view), 1: staged in LDS, 2: too large, gather from global memory
__device__ inline int window_mode(int wvn, int wun) {
    if (wvn <= 0 || wun <= 0) return 0;
    return (wvn * wun <= kWinCap) ? 1 : 2;
}

// Work split of the backprojection (kind): the fast pass (kBpFast) handles the common case
// only -- LDS window, z chunk on one side of the source (affine footprints), at most kMaxRows
// rows and kMaxUTaps columns, slowly varying footprint heights, no weight zeroing -- and
// reports every (work-item, view) it
// leaves; the completing pass (kBpRest) does exactly those; kBpAll does everything.  Keeping
// the rare cases out of the fast pass keeps its register count, and so its occupancy, low.
constexpr int kBpFast = 0, kBpAll = 1, kBpRest = 2;

// Footprint of a voxel column for one view: the four corners (+-hx, +-hy) rotated into the
// source frame give the u-footprint [ulo, uhi] (index space) and the range [rmin, rmax] of
// 1/(Sy - y') over the corners.  Every pass computes it here, so all passes round it alike.
struct BpColumn {
    float yr0, rmin, rmax, ulo, uhi;
    int ku0, ku1;
};

__device__ inline BpColumn bp_column(const BpView& v, float X, float Y, float hx, float hy, int curved) {
    BpColumn col;
    const float xr0 = fmaf(X, v.c, Y * v.s);
    col.yr0 = fmaf(Y, v.c, -(X * v.s));
    // the corners (+-hx, +-hy) rotate to (xr0, yr0) +- (P, M) and (xr0, yr0) +- (Q, -N)
    const float P = fmaf(hx, v.c, hy * v.s), Q = fmaf(hx, v.c, -(hy * v.s));
    const float M = fmaf(hy, v.c, -(hx * v.s)), N = fmaf(hy, v.c, hx * v.s);
    // corner first, then its distance from the source: one rounding at the source's scale
    const float r0 = __builtin_amdgcn_rcpf(v.Sy - (col.yr0 + M)), r1 = __builtin_amdgcn_rcpf(v.Sy - (col.yr0 - M));
    const float r2 = __builtin_amdgcn_rcpf(v.Sy - (col.yr0 - N)), r3 = __builtin_amdgcn_rcpf(v.Sy - (col.yr0 + N));
    const float su0 = (v.Sx - (xr0 + P)) * r0, su1 = (v.Sx - (xr0 - P)) * r1;
    const float su2 = (v.Sx - (xr0 + Q)) * r2, su3 = (v.Sx - (xr0 - Q)) * r3;
    const float smin = fminf(fminf(su0, su1), fminf(su2, su3));
    const float smax = fmaxf(fmaxf(su0, su1), fmaxf(su2, su3));
    col.rmin = fminf(fminf(r0, r1), fminf(r2, r3));
    col.rmax = fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
    const float fu1 = u_index(v, smin, curved), fu2 = u_index(v, smax, curved);
    col.ulo = fminf(fu1, fu2);
    col.uhi = fmaxf(fu1, fu2);
    col.ku0 = (int)floorf(col.ulo);
    col.ku1 = (int)floorf(col.uhi);
    return col;
}

// v-footprints (index space) of the ZC voxels of a chunk.  Voxel i has e = Sz - z = e0 - i dz and
// its footprint ends are g1 = Av - Bv (e - hz) r1, g2 = Av - Bv (e + hz) r2 with r1 = rmin,
// r2 = rmax where e -/+ hz >= 0 and r1 = rmax, r2 = rmin where e -/+ hz <= 0.  When the whole
// chunk lies on one side of the source height (uniform across the workgroup: `lin`) both ends
// are affine in i, lo = Alo + i Dlo and hi = Ahi + i Dhi, and g2 - g1 = -Bv (smx - smn) has the
// sign of -Bv for every voxel.
struct BpChunk {
    float Alo, Dlo, Ahi, Dhi;
    bool lin;
};

template <int ZC>
__device__ inline BpChunk bp_chunk(const BpView& v, const BpColumn& col, float Z0, float dz, float hz) {
    BpChunk ch;
    const float e0 = v.Sz - Z0;
    const float eL = fmaf(-(float)(ZC - 1), dz, e0);
    const bool above = fminf(e0, eL) - hz >= 0.0f;
    ch.lin = above || (fmaxf(e0, eL) + hz <= 0.0f);
    const float r1 = above ? col.rmin : col.rmax, r2 = above ? col.rmax : col.rmin;
    const float B1 = v.Bv * r1, B2 = v.Bv * r2;
    const float A1 = fmaf(-B1, e0 - hz, v.Av), D1 = B1 * dz;
    const float A2 = fmaf(-B2, e0 + hz, v.Av), D2 = B2 * dz;
    const bool g1_high = v.Bv > 0.0f;
    ch.Alo = g1_high ? A2 : A1;
    ch.Dlo = g1_high ? D2 : D1;
    ch.Ahi = g1_high ? A1 : A2;
    ch.Dhi = g1_high ? D1 : D2;
    return ch;
}

// Row-integral table of a work-item (its LDS column `scan`, stride 256 floats): entry m holds
// C at row boundary ks - 1 + m, so entries 0 and 1 are 0 (boundaries ks - 1 and ks), entry j + 2
// the integral of rows ks .. ks + j, and entry nr + 2 repeats the total.  A lookup at
// x = v - (ks - 1) interpolates entries floor(x) and floor(x) + 1: for x in [0, nr + 2) it needs
// no clamp, which leaves a row of slack on either side of the rows [ks, ks + nr] for rounding.
__device__ inline float bp_lookup(const float* __restrict__ scan, float x) {
    const int jj = (int)x;
    const float a = scan[jj * 256], b = scan[(jj + 1) * 256];
    return fmaf(__builtin_amdgcn_fractf(x), b - a, a);
}

// The row integral of nr window rows from `rp` (stride `stride`), u-filtered with three taps
// (weight 0 past the footprint; the window buffers are zero-initialised and padded, so such
// reads stay finite) and times W, into the table above.  The tap weights `wa` come with W
// folded in (wa[t] = wu[t] * W), so a row is three FMAs into the running sum, taps in the order
// 2, 1, 0 (every pass accumulates a row this way).  Two rows per step.
__device__ inline void bp_row_integral3(const float* __restrict__ rp, int stride, int nr, const float (&wa)[kMaxUTaps],
                                        float* __restrict__ scan) {
    const float a0 = wa[0], a1 = wa[1], a2 = wa[2];
    float csum = 0.0f;
    scan[0] = 0.0f;
    scan[256] = 0.0f;
    float* sp = scan + 256;
    // the loop runs on the row pointer (no row counter to keep per lane)
    const float* const pair_end = rp + __mul24(nr & ~1, stride);
    while (rp != pair_end) {
        const float* rq = rp + stride;
        const float p0 = rp[0], p1 = rp[1], p2 = rp[2];
        const float q0 = rq[0], q1 = rq[1], q2 = rq[2];
        csum = fmaf(a0, p0, fmaf(a1, p1, fmaf(a2, p2, csum)));
        sp[256] = csum;
        csum = fmaf(a0, q0, fmaf(a1, q1, fmaf(a2, q2, csum)));
        sp[512] = csum;
        sp += 512;
        rp += 2 * stride;
    }
    if (nr & 1) {
        csum = fmaf(a0, rp[0], fmaf(a1, rp[1], fmaf(a2, rp[2], csum)));
        sp[256] = csum;
    }
    scan[(nr + 2) * 256] = csum;
}

// The voxels of a sub-chunk (acc[OFF .. OFF+kBpSub-1]) with affine footprints from the row
// integral: C between row boundaries is linear.  With CLAMP, v - (ks - 1) is clamped to
// [1, nr + 1], which gives C = 0 below and the total above (rows outside the detector or the
// window contribute nothing).  Without it the footprints of all kBpSub voxels must lie inside
// the table's rows (bp_sub_rows' `inside`: the sub-chunk is whole and its rows were not cut by
// the detector or the window); the clamp then never binds, so both ways give the same sums.
// Without EXACT, 1/(hi - lo) of voxel i comes from a quadratic in t = i - 7.5 (the sub-chunk's
// middle): the footprint height is affine in i, h_i = h_m (1 + t e) with e the fraction it
// changes by per voxel (the magnification spread over one voxel), and
// 1/h_i = r_m (1 - t e + (t e)^2) up to a relative error below |t e|^3 <= (7.5 e)^3.  With t a
// constant of the unrolled loop that is two FMAs per voxel and no reciprocal (which issues at
// half the FMA rate).  It is used only where bp_series_ok holds for the voxel column
// (e <= kSeriesRel, error below 3.4e-6; e is ~3.5e-4 at 1000 mm from the source for 0.25 mm
// voxels); with EXACT every voxel takes its own reciprocal.
constexpr float kSeriesRel = 2e-3f;

// The per-voxel change d of the footprint height of any affine sub-chunk of a column: with
// h_0 = |Ahi - Alo| >= |Bv| hz (rmax + rmin) and |Dhi - Dlo| = |Bv| 2 hz (rmax - rmin) on either
// side of the source height, d <= 2 (rmax - rmin) / (rmax + rmin), a bound of the column alone.
__device__ inline bool bp_series_ok(const BpColumn& col) {
    return 2.0f * (col.rmax - col.rmin) <= kSeriesRel * (col.rmax + col.rmin);
}

template <int OFF, bool CLAMP, bool EXACT, int ZC>
__device__ inline void bp_voxels_affine(const BpChunk& ch, int ks, int nr, const float* __restrict__ scan,
                                        float (&acc)[ZC]) {
    static_assert(OFF + kBpSub <= ZC, "sub-chunk inside the work-item");
    const float base = (float)(ks - 1), top = (float)(nr + 1);
    const float alo = ch.Alo - base, ahi = ch.Ahi - base;
    // 1/h_i = rm + t c1 + t^2 c2 with t = i - 7.5, rm = 1/h(7.5), e = (Dhi - Dlo) rm,
    // c1 = -rm e, c2 = rm e^2
    constexpr float kMid = 0.5f * (float)(kBpSub - 1);
    const float dD = ch.Dhi - ch.Dlo;
    const float rm = EXACT ? 0.0f : __builtin_amdgcn_rcpf(fmaf(kMid, dD, ahi - alo));
    const float e = dD * rm;
    const float c1 = -(rm * e), c2 = -(c1 * e);
#pragma unroll
    for (int i = 0; i < kBpSub; ++i) {
        const float lo = fmaf((float)i, ch.Dlo, alo), hi = fmaf((float)i, ch.Dhi, ahi);
        const float t = (float)i - kMid;
        float rinv;
        if (EXACT)
            rinv = __builtin_amdgcn_rcpf(hi - lo);
        else
            rinv = fmaf(t * t, c2, fmaf(t, c1, rm));
        const float xl = CLAMP ? __builtin_amdgcn_fmed3f(lo, 1.0f, top) : lo;
        const float xh = CLAMP ? __builtin_amdgcn_fmed3f(hi, 1.0f, top) : hi;
        acc[OFF + i] = fmaf(bp_lookup(scan, xh) - bp_lookup(scan, xl), rinv, acc[OFF + i]);
    }
}

// z of the first voxel of sub-chunk h of a work-item whose first voxel is at Z0: every pass
// derives its sub-chunks from this, so they round them alike
__device__ inline float bp_sub_z0(float Z0, float dz, int h) { return fmaf((float)(kBpSub * h), dz, Z0); }

// Rows [ks, ks + nr - 1] (clamped to the detector and to the rows [row0, row0 + nrows) of the
// source) that the n voxels of an affine sub-chunk touch: the ends move monotonically with i,
// so the range comes from the end voxels.  `inside`: the sub-chunk is whole and nothing was cut.
__device__ inline void bp_sub_rows(const BpChunk& ch, int n, int row0, int nrows, int nv, int& ks, int& nr,
                                   bool& inside) {
    const float fL = (float)(n - 1);
    const float lo0 = fmaf(0.0f, ch.Dlo, ch.Alo), hi0 = fmaf(0.0f, ch.Dhi, ch.Ahi);
    const float loL = fmaf(fL, ch.Dlo, ch.Alo), hiL = fmaf(fL, ch.Dhi, ch.Ahi);
    const int klo = (int)floorf(fminf(lo0, loL)), khi = (int)floorf(fmaxf(hi0, hiL));
    ks = max(max(klo, 0), row0);
    const int ke = min(min(khi, nv - 1), row0 + nrows - 1);
    nr = ke - ks + 1;
    inside = n == kBpSub && ks == klo && ke == khi;
}

// Calls f(integral_constant<int, h>) for every sub-chunk h of a ZC-voxel work-item.
template <int ZC, int H = 0, typename F>
__device__ inline void bp_for_subs(F&& f) {
    if constexpr (H * kBpSub < ZC) {
        f(std::integral_constant<int, H>{});
        bp_for_subs<ZC, H + 1>(f);
    }
}

// The fast pass (kBpFast): the common case only -- window in LDS, sub-chunk on one side of the
// source height, at most kMaxRows rows and kMaxUTaps columns, a footprint height that changes
// slowly enough along z for the series normaliser (bp_series_ok), no weight zeroing (the
// host runs this pass only with zero_tol = 0).  The column's u-footprint, u-weights and FDK weight are
// worked out once per view for all ZC voxels; every sub-chunk then builds its own row integral.
// Returns true when it leaves a sub-chunk of this (work-item, view) to the completing pass,
// which decides from the same bp_column / bp_chunk values.
template <int ZC>
__device__ inline bool bp_column_view_fast(const BpView& v, const float* __restrict__ win, int stride,
                                           int row0, int nrows, int col0, int ncols, float X, float Y,
                                           float Z0, float dz, float hx, float hy, float hz, int nz_left,
                                           int curved, float scale, int nu, int nv, float* __restrict__ scan,
                                           float (&acc)[ZC]) {
    const BpColumn col = bp_column(v, X, Y, hx, hy, curved);
    // taps inside the detector and the window
    const int kus = max(max(col.ku0, 0), col0), kue = min(min(col.ku1, nu - 1), col0 + ncols - 1);
    if (kus > kue) return false;
    const bool taps_ok = kue - kus + 1 <= kMaxUTaps;
    // overlap of each tap with [ulo, uhi] over the footprint width (python/GPUFuncs_pycuda.py:
    // 427-438 without zeroing); a footprint inside one cell weighs 1
    const float inv_u = __builtin_amdgcn_rcpf(col.uhi - col.ulo);
    const float fk = (float)kus;
    float wu[kMaxUTaps];
#pragma unroll
    for (int j = 0; j < kMaxUTaps; ++j) {
        const float k = fk + (float)j;
        const float w = (fminf(col.uhi, k + 1.0f) - fmaxf(col.ulo, k)) * inv_u;
        wu[j] = (kus + j <= kue) ? fmaxf(w, 0.0f) : 0.0f;
    }
    if (col.ku0 == col.ku1) wu[0] = 1.0f;
    const bool series_ok = bp_series_ok(col);
    const float dd = v.R - col.yr0;
    const float W = scale * (v.R * v.R) * __builtin_amdgcn_rcpf(dd * dd);
    float wa[kMaxUTaps];
#pragma unroll
    for (int j = 0; j < kMaxUTaps; ++j) wa[j] = wu[j] * W;
    const float* wp = win + (kus - col0);
    bool left = false;
    bp_for_subs<ZC>([&](auto hc) {
        constexpr int h = decltype(hc)::value;
        const int n = min(kBpSub, nz_left - kBpSub * h);
        if (n <= 0) return;
        const BpChunk ch = bp_chunk<kBpSub>(v, col, bp_sub_z0(Z0, dz, h), dz, hz);
        if (!ch.lin) {
            left = true;
            return;
        }
        int ks, nr;
        bool inside;
        bp_sub_rows(ch, n, row0, nrows, nv, ks, nr, inside);
        if (nr <= 0) return;
        if (nr > kMaxRows || !taps_ok || !inside || !series_ok) {
            left = true;
            return;
        }
        // window offsets fit 24 bits: a full-rate multiply (the 32-bit one is quarter rate)
        bp_row_integral3(wp + __mul24(ks - row0, stride), stride, nr, wa, scan);
        bp_voxels_affine<kBpSub * h, false, false>(ch, ks, nr, scan, acc);
    });
    return left;
}

// The completing pass (kBpRest: every (work-item, view, sub-chunk) the fast pass left, and
// nothing else) and the single pass over all cases (kBpAll).
template <int ZC, bool FROM_LDS, int KIND>
__device__ inline void bp_column_view(const BpView& v, const float* __restrict__ src, int stride,
                                      int row0, int nrows, int col0, int ncols, float X, float Y,
                                      float Z0, float dz, float hx, float hy, float hz, int nz_left,
                                      int curved, float scale, float zero_tol, int nu, int nv,
                                      bool window_in_lds, float* __restrict__ scan, float (&acc)[ZC]) {
    static_assert(KIND == kBpAll || KIND == kBpRest, "the fast pass has its own function");
    const BpColumn col = bp_column(v, X, Y, hx, hy, curved);
    const float ulo = col.ulo, uhi = col.uhi;
    const int ku0 = col.ku0, ku1 = col.ku1;
    // taps inside the detector and the source window
    const int kus = max(max(ku0, 0), col0), kue = min(min(ku1, nu - 1), col0 + ncols - 1);
    if (kus > kue) return;
    const float inv_u = __builtin_amdgcn_rcpf(uhi - ulo);
    float wu[kMaxUTaps];
#pragma unroll
    for (int j = 0; j < kMaxUTaps; ++j) {
        const int ku = kus + j;
        wu[j] = (ku <= kue) ? dd_weight(ku, ku0, ku1, ulo, uhi, inv_u, zero_tol) : 0.0f;
    }
    const float dd = v.R - col.yr0;
    const float W = scale * (v.R * v.R) * __builtin_amdgcn_rcpf(dd * dd);
    float wa[kMaxUTaps];
#pragma unroll
    for (int j = 0; j < kMaxUTaps; ++j) wa[j] = wu[j] * W;
    const float rmin = col.rmin, rmax = col.rmax;
    const int cbase = kus - col0;
    const int ntaps = kue - kus + 1;
    bp_for_subs<ZC>([&](auto hc) {
        constexpr int h = decltype(hc)::value;
        constexpr int OFF = kBpSub * h;
        const int n = min(kBpSub, nz_left - kBpSub * h);
        if (n <= 0) return;
        const float Z0h = bp_sub_z0(Z0, dz, h);
        const BpChunk ch = bp_chunk<kBpSub>(v, col, Z0h, dz, hz);
        const bool lin = ch.lin;
        auto foot = [&](float fi, float& lo, float& hi) {
            if (lin) {
                lo = fmaf(fi, ch.Dlo, ch.Alo);
                hi = fmaf(fi, ch.Dhi, ch.Ahi);
            } else {
                const float e = v.Sz - fmaf(fi, dz, Z0h);
                const float ep = e + hz, em = e - hz;
                const float smx = ep * (ep >= 0.0f ? rmax : rmin);
                const float smn = em * (em >= 0.0f ? rmin : rmax);
                const float g1 = fmaf(-v.Bv, smn, v.Av), g2 = fmaf(-v.Bv, smx, v.Av);
                lo = fminf(g1, g2);
                hi = fmaxf(g1, g2);
            }
        };
        int ks, nr;
        bool inside = false;
        if (lin) {
            bp_sub_rows(ch, n, row0, nrows, nv, ks, nr, inside);
        } else {
            // the ends move monotonically with i, so the range comes from the end voxels
            float lo0, hi0, loL, hiL;
            foot(0.0f, lo0, hi0);
            foot((float)(n - 1), loL, hiL);
            const float rlo = fminf(lo0, loL), rhi = fmaxf(hi0, hiL);
            ks = max(max((int)floorf(rlo), 0), row0);
            const int ke = min(min((int)floorf(rhi), nv - 1), row0 + nrows - 1);
            nr = ke - ks + 1;
        }
        if (nr <= 0) return;
        if constexpr (KIND == kBpRest) {
            // the sub-chunks the fast pass did
            const bool common = window_in_lds && lin && inside && nr <= kMaxRows && ntaps <= kMaxUTaps &&
                                zero_tol <= 0.0f && bp_series_ok(col);
            if (common) return;
        }
        if (zero_tol <= 0.0f && nr <= kMaxRows) {
            const float* rp = src + (FROM_LDS ? __mul24(ks - row0, stride) : (size_t)(ks - row0) * stride) + cbase;
            if (FROM_LDS && !__any(ntaps > kMaxUTaps)) {
                bp_row_integral3(rp, stride, nr, wa, scan);
            } else {
                // taps from the last to the first, each weight times W, as bp_row_integral3 does
                // (the same sums for at most three taps)
                float csum = 0.0f;
                scan[0] = 0.0f;
                scan[256] = 0.0f;
                for (int j = 0; j < nr; ++j) {
                    for (int t = ntaps - 1; t >= kMaxUTaps; --t)
                        csum = fmaf(dd_weight(kus + t, ku0, ku1, ulo, uhi, inv_u, zero_tol) * W, rp[t], csum);
#pragma unroll
                    for (int t = kMaxUTaps - 1; t >= 0; --t) csum = fmaf(wa[t], rp[min(t, ntaps - 1)], csum);
                    scan[(j + 2) * 256] = csum;
                    rp += stride;
                }
                scan[(nr + 2) * 256] = csum;
            }
            if (lin) {
                bp_voxels_affine<OFF, true, true>(ch, ks, nr, scan, acc);
            } else {
                const float base = (float)(ks - 1), top = (float)(nr + 1);
#pragma unroll
                for (int i = 0; i < kBpSub; ++i) {
                    float lo, hi;
                    foot((float)i, lo, hi);
                    lo -= base;
                    hi -= base;
                    const float xl = __builtin_amdgcn_fmed3f(lo, 1.0f, top), xh = __builtin_amdgcn_fmed3f(hi, 1.0f, top);
                    acc[OFF + i] = fmaf(bp_lookup(scan, xh) - bp_lookup(scan, xl), __builtin_amdgcn_rcpf(hi - lo),
                                        acc[OFF + i]);
                }
            }
            return;
        }
        // explicit per-row weights (zeroing of small weights, or a sub-chunk over more than
        // kMaxRows rows): rare, so the voxel footprints are recomputed per row rather than held
        // in registers
        const int ke = ks + nr - 1;
        for (int k = ks; k <= ke; ++k) {
            const float* row = src + (size_t)(k - row0) * stride + cbase;
            float rk = 0.0f;
#pragma unroll
            for (int j = 0; j < kMaxUTaps; ++j) rk = fmaf(wu[j], row[min(j, ntaps - 1)], rk);
            for (int j = kMaxUTaps; j < ntaps; ++j)
                rk = fmaf(dd_weight(kus + j, ku0, ku1, ulo, uhi, inv_u, zero_tol), row[j], rk);
            rk *= W;
            const float fk = (float)k, fk1 = (float)(k + 1);
#pragma unroll
            for (int i = 0; i < kBpSub; ++i) {
                float vlo, vhi;
                foot((float)i, vlo, vhi);
                float w = (fminf(vhi, fk1) - fmaxf(vlo, fk)) * __builtin_amdgcn_rcpf(vhi - vlo);
                w = (w < zero_tol || w < 0.0f) ? 0.0f : w;
                acc[OFF + i] = fmaf(w, rk, acc[OFF + i]);
            }
        }
    });
}

// KIND: kBpFast writes the volume and sets flags[tile][view] for every view it leaves (in
// part or whole) to kBpRest, which adds those contributions; kBpAll does everything.
template <int ZC, int KIND>
__global__ void __launch_bounds__(256) dd_backproject_tile_kernel(
    const float* __restrict__ q, float* __restrict__ vol, const float* __restrict__ vt,
    uint8_t* __restrict__ flags, int nx, int ny, int nz, int nu, int nv, int n_views, float x0,
    float y0, float z0, float dx, float dy, float dz, int curved, float scale, float zero_tol,
    int accumulate, int tiles_x, int tiles_y, int tiles_z) {
    __shared__ float win[2][kWinCap + kMaxUTaps - 1];
    __shared__ int4 wring[kWinRing];
    __shared__ float scan_lds[(kMaxRows + 3) * 256];
    float* scan = scan_lds + threadIdx.x;
    // XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a contiguous range
    const int T = tiles_x * tiles_y * tiles_z;
    const int b = blockIdx.x;
    const int xcd = b & 7, local = b >> 3, qd = T >> 3, rm = T & 7;
    const int t = (xcd < rm) ? xcd * (qd + 1) + local : rm * (qd + 1) + (xcd - rm) * qd + local;
    const int tx = t % tiles_x;
    const int ty = (t / tiles_x) % tiles_y;
    const int tz = t / (tiles_x * tiles_y);
    uint8_t* __restrict__ tflags = (KIND == kBpAll) ? nullptr : flags + (size_t)t * n_views;
    if constexpr (KIND == kBpRest) {
        // nothing left for this tile: done
        int any = 0;
        for (int vw = threadIdx.x; vw < n_views; vw += 256) any |= tflags[vw];
        if (!__syncthreads_or(any)) return;
    }
    const int lx = threadIdx.x & (kTileX - 1), ly = threadIdx.x / kTileX;
    const int ix = tx * kTileX + lx, iy = ty * kTileY + ly, iz0 = tz * ZC;
    const bool active = (ix < nx) && (iy < ny);
    const float hx = 0.5f * fabsf(dx), hy = 0.5f * fabsf(dy), hz = 0.5f * fabsf(dz);
    const float X = fmaf((float)ix, dx, x0);
    const float Y = fmaf((float)iy, dy, y0);
    const float Z0 = fmaf((float)iz0, dz, z0);
    const int nz_left = min(ZC, nz - iz0);
    // tile bounding box
    const float xa0 = fmaf((float)(tx * kTileX), dx, x0);
    const float xb0 = fmaf((float)min(tx * kTileX + kTileX - 1, nx - 1), dx, x0);
    const float ya0 = fmaf((float)(ty * kTileY), dy, y0);
    const float yb0 = fmaf((float)min(ty * kTileY + kTileY - 1, ny - 1), dy, y0);
    const float za0 = Z0;
    const float zb0 = fmaf((float)(iz0 + nz_left - 1), dz, z0);
    const float xa = fminf(xa0, xb0) - hx, xb = fmaxf(xa0, xb0) + hx;
    const float ya = fminf(ya0, yb0) - hy, yb = fmaxf(ya0, yb0) + hy;
    const float za = fminf(za0, zb0) - hz, zb = fmaxf(za0, zb0) + hz;

    float acc[ZC];
#pragma unroll
    for (int i = 0; i < ZC; ++i) acc[i] = 0.0f;

    // Views are taken in pairs: the windows of views v+2 and v+3 are loaded into registers while
    // views v and v+1 are computed from the group's two LDS windows, then stored into them between
    // two barriers, so a window load has two views of work to arrive.  Two windows of 1024 cells,
    // a 64-view ring and 22 table rows per work-item keep the group at 31 KB of LDS, so five
    // groups share a CU: one more table row (32 KB) leaves room for four, 7 % slower at
    // 512^3 x 720 views.
    // The windows depend on the tile and the view only: they are worked out cooperatively, one
    // view per thread, kWinRing/2 views ahead, into a ring in LDS that every thread reads.
    auto fill_windows = [&](int first, int count) {
        for (int t = threadIdx.x; t < count; t += 256) {
            const int view = first + t;
            int4 w = make_int4(0, 0, 0, 0);
            if (view < n_views)
                tile_window(load_view(vt, view), xa, xb, ya, yb, za, zb, curved, nu, nv, w.x, w.y, w.z, w.w);
            wring[view & (kWinRing - 1)] = w;
        }
    };
    // every thread reads the same ring entry: the window is wave-uniform (scalar registers)
    auto window_of = [&](int view, int& v0, int& vn, int& u0, int& un) {
        const int4 w = wring[view & (kWinRing - 1)];
        v0 = __builtin_amdgcn_readfirstlane(w.x);
        vn = __builtin_amdgcn_readfirstlane(w.y);
        u0 = __builtin_amdgcn_readfirstlane(w.z);
        un = __builtin_amdgcn_readfirstlane(w.w);
    };
    fill_windows(0, kWinRing);
    __syncthreads();
    auto fetch = [&](int view, int v0, int vn, int u0, int un, float (&pf)[kPrefetch]) {
        if (window_mode(vn, un) != 1) return;
        if constexpr (KIND == kBpRest) {  // only the views the fast pass left
            if (view >= n_views || !tflags[view]) return;
        }
        const float* qn = q + (size_t)view * nv * nu + (size_t)v0 * nu + u0;
        // element e = threadIdx.x + 256 p of the window is (row r, column c); step by 256.
        // floor((e + 1/2) / un) in FP32: the quotient is at least 0.5/un from an integer and
        // its rounding error below 3e-5/un, so the floor is exact (e <= 256, un <= kWinCap)
        const float inv_un = __builtin_amdgcn_rcpf((float)un);
        int r = (int)(((float)threadIdx.x + 0.5f) * inv_un);
        int c = (int)threadIdx.x - __mul24(r, un);
        const int dr = (int)(256.5f * inv_un), dc = 256 - __mul24(dr, un);
#pragma unroll
        for (int p = 0; p < kPrefetch; ++p) {
            if (r < vn) pf[p] = qn[__mul24(r, nu) + c];
            r += dr;
            c += dc;
            if (c >= un) {
                c -= un;
                ++r;
            }
        }
    };
    auto put = [&](float* buf, int view, int vn, int un, const float (&pf)[kPrefetch]) {
        if (window_mode(vn, un) != 1) return;
        if constexpr (KIND == kBpRest) {
            if (view >= n_views || !tflags[view]) return;
        }
#pragma unroll
        for (int p = 0; p < kPrefetch; ++p) {
            const int e = threadIdx.x + p * 256;
            if (e < vn * un) buf[e] = pf[p];
        }
    };
    // zero every window buffer once: cells past a window are read with weight 0
    for (int e = threadIdx.x; e < (int)(sizeof(win) / sizeof(float)); e += 256) (&win[0][0])[e] = 0.0f;
    __syncthreads();
    // windows (first row, rows, first column, columns) of the pair in the current buffers
    int w0[2], wn[2], wu[2], wun[2];
    for (int k = 0; k < 2; ++k) {
        window_of(k, w0[k], wn[k], wu[k], wun[k]);
        if (window_mode(wn[k], wun[k]) == 1) {
            for (int e = threadIdx.x; e < wn[k] * wun[k]; e += 256)
                win[k][e] = q[(size_t)k * nv * nu + (size_t)(w0[k] + e / wun[k]) * nu + wu[k] + e % wun[k]];
        }
    }
    __syncthreads();
    for (int view = 0; view < n_views; view += 2) {
        // refill the half of the ring last used by views view - kWinRing/2 .. view - 1 (read
        // before the previous barrier) with views view + kWinRing/2 .. view + kWinRing - 1
        if (view > 0 && (view & (kWinRing / 2 - 1)) == 0) fill_windows(view + kWinRing / 2, kWinRing / 2);
        int x0n[2], xnn[2], xun[2], xunn[2];  // windows of views v+2, v+3
        float pf0[kPrefetch], pf1[kPrefetch];
        window_of(view + 2, x0n[0], xnn[0], xun[0], xunn[0]);
        window_of(view + 3, x0n[1], xnn[1], xun[1], xunn[1]);
        fetch(view + 2, x0n[0], xnn[0], xun[0], xunn[0], pf0);
        fetch(view + 3, x0n[1], xnn[1], xun[1], xunn[1], pf1);
#pragma unroll 1
        for (int k = 0; k < 2; ++k) {
            const int vw = view + k;
            const int mode = window_mode(wn[k], wun[k]);
            if (vw >= n_views || mode == 0) continue;
            if constexpr (KIND == kBpRest) {
                if (!tflags[vw]) continue;
            }
            if constexpr (KIND == kBpFast) {
                if (mode != 1) {  // window too large for LDS: the whole view is left
                    if (threadIdx.x == 0) tflags[vw] = 1;
                    continue;
                }
                if (!active) continue;
                if (bp_column_view_fast<ZC>(load_view(vt, vw), win[k], wun[k], w0[k], wn[k], wu[k], wun[k], X, Y,
                                            Z0, dz, hx, hy, hz, nz_left, curved, scale, nu, nv, scan, acc))
                    tflags[vw] = 1;
            } else {
                if (!active) continue;
                const BpView cur = load_view(vt, vw);
                if (mode == 1) {
                    bp_column_view<ZC, true, KIND>(cur, win[k], wun[k], w0[k], wn[k], wu[k], wun[k], X, Y,
                                                   Z0, dz, hx, hy, hz, nz_left, curved, scale, zero_tol, nu, nv,
                                                   true, scan, acc);
                } else {
                    bp_column_view<ZC, false, KIND>(cur, q + (size_t)vw * nv * nu, nu, 0, nv, 0, nu, X, Y, Z0,
                                                    dz, hx, hy, hz, nz_left, curved, scale, zero_tol, nu, nv,
                                                    false, scan, acc);
                }
            }
        }
        __syncthreads();  // the pair's windows are read
        put(win[0], view + 2, xnn[0], xunn[0], pf0);
        put(win[1], view + 3, xnn[1], xunn[1], pf1);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            w0[k] = x0n[k]; wn[k] = xnn[k]; wu[k] = xun[k]; wun[k] = xunn[k];
        }
    }
    if (!active) return;
#pragma unroll
    for (int i = 0; i < ZC; ++i) {
        if (i >= nz_left) break;
        const size_t idx = ((size_t)(iz0 + i) * ny + iy) * nx + ix;
        if constexpr (KIND == kBpRest)
            vol[idx] += acc[i];
        else
            vol[idx] = accumulate ? vol[idx] + acc[i] : acc[i];
    }
}

__global__ void __launch_bounds__(256) fill_zero_kernel(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ------------------------------------------------------------------------------------------
// Matched adjoint of the forward projector
// ------------------------------------------------------------------------------------------

// The exact transpose of dd_forward_kernel, as a gather over voxels (no atomics).  No reference
// counterpart: the reference backprojects with the FDK-weighted operator, which is not A^T.
// For one view a voxel sits in slab n at transverse cell ka and z cell kc (index space of the
// forward projector's tables).  The detector columns whose U-edge lines bracket [ka, ka+1] at
// slab n are found by inverting the central projection through the source (edge coordinate
// of the ray through a point), with a rounding margin; every candidate then gets the
// forward projector's own transverse weight o_a / w_a (overlaps below 1e-6 of the footprint
// dropped, as there).
//
// z direction: the forward projector gives row k of column j at this slab
// L_k o_a / (w_a tt (nu_k+1 - nu_k)) times the z integral of the voxel line over the row's
// footprint [c_k, c_k+1], c_e = nu_e tt + gamma, with L_k = sec_j sqrt(1 + s2_k^2).  The row
// edges are equally spaced in nu (nu_e = nuA + nuB e), so a row's footprint is the unit step
// [k, k+1] in the edge coordinate e(c) = ((c - gamma) / tt - nuA) / nuB, and voxel kc receives
// sgn(nuB tt) o_a / w_a (Y(e(kc + 1)) - Y(e(kc))) with Y the linear interpolation of the
// row prefix Y_e = sum_{k < e} y_k L_k (adj_rowprefix_kernel): one table lookup per voxel
// boundary instead of a walk over rows.
constexpr int kAdjViewStride = 16;
#ifndef CTR_ADJ_ZC
#define CTR_ADJ_ZC 16
#endif
constexpr int kAdjZChunk = CTR_ADJ_ZC;  // voxels along z per work-item

// Per-view constants (FP64 -> FP32):
//   [0] branch  [1] Pt: source transverse index coordinate  [2] Ps: source slab coordinate
//   [3] r = d_t / d_s  [4] eu_t [5] eu_s [6] ew_t [7] ew_s  (detector axes, transverse/slab parts)
//   [8] SDD  [9] u2_0: coordinate of edge 0  [10] 1/step (mm or rad)
//   [11] nuA, [12] 1/nuB : row edge e has z index slope nu_e = nuA + nuB * e
__global__ void adj_geometry_kernel(ctr_geometry g, const double* __restrict__ angles,
                                    const int32_t* __restrict__ branch, float* __restrict__ at) {
    const int view = blockIdx.x * blockDim.x + threadIdx.x;
    if (view >= g.n_views) return;
    const double th = angles[view];
    double s, c;
    sincos(th, &s, &c);
    const double Sx = -g.sad * s, Sy = g.sad * c;
    const int br = branch[view];
    const bool xs = (br == CTR_BRANCH_X);
    const double ds = xs ? g.dx : -g.dy;
    const double dt = xs ? -g.dy : g.dx;
    const int ns = xs ? g.nx : g.ny;
    const int nt = xs ? g.ny : g.nx;
    const double rcs = xs ? g.recon_center[0] : g.recon_center[1];
    const double rct = xs ? g.recon_center[1] : g.recon_center[0];
    const double Ss = xs ? Sx : Sy, St = xs ? Sy : Sx;
    const double s0 = rcs - (ns - 1) / 2.0 * ds;
    // eu = (cos, sin), ew = (sin, -cos) in (x, y)
    const double eux = c, euy = s, ewx = s, ewy = -c;
    const double step = (g.detector_shape == CTR_DETECTOR_FLAT) ? g.du : g.du / g.sdd;
    float* T = at + (size_t)view * kAdjViewStride;
    T[0] = (float)br;
    T[1] = (float)((St - rct) / dt + nt / 2.0);
    T[2] = (float)((Ss - s0) / ds);
    T[3] = (float)(dt / ds);
    T[4] = (float)(xs ? euy : eux);
    T[5] = (float)(xs ? eux : euy);
    T[6] = (float)(xs ? ewy : ewx);
    T[7] = (float)(xs ? ewx : ewy);
    T[8] = (float)g.sdd;
    T[9] = (float)(-(g.nu - 1) / 2.0 * step - step / 2.0 + g.det_offset[0]);
    T[10] = (float)(1.0 / step);
    // row edges: v2_e = -(e - (nv-1)/2) dv + half + off, nu_e = v2_e / (-dz)
    const double half = (g.detector_shape == CTR_DETECTOR_FLAT) ? 0.5 * g.dv : -0.5 * g.dv;
    const double nuA = ((g.nv - 1) / 2.0 * g.dv + half + g.det_offset[1]) / (-g.dz);
    const double nuB = (-g.dv) / (-g.dz);
    T[11] = (float)nuA;
    T[12] = (float)(1.0 / nuB);
}

// Edge coordinate (fractional edge index) of the ray from the source through the point at
// transverse index coordinate tk and slab coordinate n.
__device__ inline float adj_edge_coord(const float* A, float tk, float n, int curved) {
    const float m = (tk - A[1]) * A[3] * __builtin_amdgcn_rcpf(n - A[2]);
    const float num = A[6] - m * A[7];
    const float den = m * A[5] - A[4];
    const float u2 = curved ? atanf(num * __builtin_amdgcn_rcpf(den)) : A[8] * num * __builtin_amdgcn_rcpf(den);
    return (u2 - A[9]) * A[10];
}

// y[view][k][j] -> Y[view][j][e], e = 0 .. nvp-1: Y_e = sum_{k < e} y_k L_k (L_k = sec_j
// sqrt(1 + s2_k^2), evaluated as the forward projector evaluates it), summed in FP64, stored in
// FP32; entries past nv repeat the total.  One workgroup per (view, 32 columns): 32 x 32 tiles
// through LDS, coalesced along j on read and along e on write; lanes 0..31 carry the sums.
__global__ void __launch_bounds__(256) adj_rowprefix_kernel(const float* __restrict__ y,
                                                            float* __restrict__ Y,
                                                            const float* __restrict__ tabs,
                                                            const float* __restrict__ rows, int nu,
                                                            int nv, int nvp, int stride) {
    __shared__ float tile[32][33];  // [k][j]
    __shared__ float line[32][33];  // [j][e]
    const int j0 = blockIdx.x * 32, view = blockIdx.y;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const float* T = tabs + (size_t)view * stride;
    const int jj = min(j0 + tx, nu - 1);
    const float sec = T[4 * nu + 2 + jj], inv = T[5 * nu + 2 + jj];
    double run = 0.0;  // running sum of column j0 + threadIdx.x (threads 0..31)
    for (int k0 = 0; k0 < nvp; k0 += 32) {
        for (int r = ty; r < 32; r += 8) {
            const int k = k0 + r, j = j0 + tx;
            float v = 0.0f;
            if (k < nv && j < nu) {
                const float s2 = rows[nv + 1 + k] * inv;
                v = y[((size_t)view * nv + k) * nu + j] * (sec * sqrtf(1.0f + s2 * s2));
            }
            tile[r][tx] = v;
        }
        __syncthreads();
        if (threadIdx.x < 32) {
            for (int r = 0; r < 32; ++r) {
                line[threadIdx.x][r] = (float)run;
                run += (double)tile[r][threadIdx.x];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int jl = e >> 5, el = e & 31;
            const int j = j0 + jl, k = k0 + el;
            if (j < nu && k < nvp) Y[((size_t)view * nu + j) * nvp + k] = line[jl][el];
        }
        __syncthreads();
    }
}

// One work-item per voxel column chunk of kAdjZChunk voxels.  Per view and candidate detector
// column the rows that reach the chunk are walked once; each row's z footprint (the forward
// projector's own fma of the same table values) is overlapped with every voxel of the chunk.
__global__ void __launch_bounds__(256) dd_adjoint_kernel(
    const float* __restrict__ Y, float* __restrict__ vol, const float* __restrict__ tabs,
    const float* __restrict__ at, int nx, int ny, int nz, int nu, int nv, int nvp, int n_views,
    int stride, int curved, float scale, float zero_tol, int accumulate) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int iz0 = blockIdx.z * kAdjZChunk;
    if (ix >= nx || iy >= ny) return;
    const int nzc = min(kAdjZChunk, nz - iz0);
    float acc[kAdjZChunk];
#pragma unroll
    for (int i = 0; i < kAdjZChunk; ++i) acc[i] = 0.0f;
    const float tol = zero_tol;
    const float fz0 = (float)iz0, fnv = (float)nv;
    for (int view = 0; view < n_views; ++view) {
        const float* A = at + (size_t)view * kAdjViewStride;
        const int br = (int)A[0];
        if (br != CTR_BRANCH_X && br != CTR_BRANCH_Y) continue;
        const bool xs = (br == CTR_BRANCH_X);
        const int n = xs ? ix : iy;     // slab
        const int ka = xs ? iy : ix;    // transverse cell
        const float fn = (float)n, fka = (float)ka;
        const float e1 = adj_edge_coord(A, fka, fn, curved);
        const float e2 = adj_edge_coord(A, fka + 1.0f, fn, curved);
        // the inverse projection is accurate to ~1e-5 edge; a 0.01 margin covers its rounding
        const int j0 = max((int)floorf(fminf(e1, e2) - 0.01f), 0);
        const int j1 = min((int)floorf(fmaxf(e1, e2) + 0.01f), nu - 1);
        const float* T = tabs + (size_t)view * stride;
        const float gam = T[6 * nu + 2];
        const float nuA = A[11], inuB = A[12];
        const float* Yv = Y + (size_t)view * nu * nvp;
        for (int j = j0; j <= j1; ++j) {
            const float a1 = fmaf(T[nu + 1 + j], fn, T[j]);
            const float a2 = fmaf(T[nu + 2 + j], fn, T[j + 1]);
            const float alo = fminf(a1, a2), ahi = fmaxf(a1, a2);
            const float wa = ahi - alo;
            const float oa = fminf(ahi, fka + 1.0f) - fmaxf(alo, fka);
            if (!(oa > 0.0f) || oa < tol * wa) continue;
            const float t = fmaf(T[3 * nu + 2 + j], fn, T[2 * nu + 2 + j]);
            const float rt = __builtin_amdgcn_rcpf(t);
            // edge coordinate of the chunk's voxel boundaries: e_i = e_0 + i de
            const float de = rt * inuB;
            const float eb = fmaf(fz0 - gam, de, -nuA * inuB);
            const float kap = (de > 0.0f ? oa : -oa) * __builtin_amdgcn_rcpf(wa);  // sgn(nuB tt) o_a / w_a
            const char* col = reinterpret_cast<const char*>(Yv + (size_t)j * nvp);
            auto lookup = [&](float e) {
                const float ec = __builtin_amdgcn_fmed3f(e, 0.0f, fnv);
                const int k = (int)ec;
                const PTaps p = *reinterpret_cast<const PTaps*>(col + 4 * k);
                return fmaf(__builtin_amdgcn_fractf(ec), p.b - p.a, p.a);
            };
            float prev = lookup(eb);
#pragma unroll
            for (int i = 0; i < kAdjZChunk; ++i) {
                const float cur = lookup(fmaf((float)(i + 1), de, eb));
                acc[i] = fmaf(kap, cur - prev, acc[i]);
                prev = cur;
            }
        }
    }
    for (int i = 0; i < nzc; ++i) {
        const size_t idx = ((size_t)(iz0 + i) * ny + iy) * nx + ix;
        const float v = acc[i] * scale;
        vol[idx] = accumulate ? vol[idx] + v : v;
    }
}

// ------------------------------------------------------------------------------------------
// FDK filter
// ------------------------------------------------------------------------------------------

constexpr int kMaxLog2N = 13;  // FFT length <= 8192 (nu <= 4097)

int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return ((1 << l) == n) ? l : -1;
}

// Filter magnitude of Reconstruction.Filter (python/Reconstruction_pycuda.py:186-221) in FFT
// order, FP64.  The reference takes |fftshift(fft(h))| of the discrete Ram-Lak taps; h is
// even, so |DFT|(w) = |h0 + 2 sum_{odd n} h(n) cos(n w)| with w = 2 pi f / N (signed f).
// The cosine table lives in LDS; arguments are reduced exactly as (f*n mod N).
// ConeBeam.Filter (python/ConeBeamGPU.py:355-384) is the same series times gain = 2 with the
// mask |w| > mask_w = pi*cutoff; Reconstruction.Filter has gain 1, mask_w = pi*cutoff/(2 tau).
__global__ void __launch_bounds__(256) spectrum_kernel(double* __restrict__ out, int N,
                                                       double tau, int ftype, double cutoff,
                                                       double gain, double mask_w) {
    extern __shared__ double costab[];
    for (int m = threadIdx.x; m < N; m += blockDim.x) costab[m] = cos(2.0 * kPi * m / N);
    __syncthreads();
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= N) return;
    const int fs = (f < N / 2) ? f : f - N;  // signed frequency; f = N/2 is -N/2 as fftshift has it
    const double h0 = 1.0 / (8.0 * tau * tau);
    double acc = 0.0;
    const int fm = ((fs % N) + N) % N;
    for (int n = 1; n < N / 2; n += 2) {
        const double hn = -0.5 / ((kPi * tau * n) * (kPi * tau * n));
        acc += hn * costab[(int)(((long long)fm * n) % N)];
    }
    double H = fabs(h0 + 2.0 * acc) * gain;
    const double w = 2.0 * kPi * fs / N;
    switch (ftype) {
        case CTR_FILTER_SHEPP_LOGAN: {
            const double a = w / (2.0 * cutoff);
            H = (fs == 0) ? 0.0 : H * sin(a) / a;  // DC: H * sin(0) = 0, as the reference
            break;
        }
        case CTR_FILTER_COSINE: H *= cos(w / (2.0 * cutoff)); break;
        case CTR_FILTER_HAMMING: H *= 0.54 + 0.46 * cos(w / cutoff); break;
        case CTR_FILTER_HANN: H *= 0.5 + 0.5 * cos(w / cutoff); break;
        default: break;
    }
    if (fabs(w) > mask_w) H = 0.0;
    out[f] = H;
}

// Twiddles of the Stockham stages of a length-N FFT, stage-major: the stage with sub-transform
// length ns (1, 16, 256, ... for the radix-16 stages, then the last stage of radix R = 2, 4 or 8)
// keeps exp(-2 pi i r m / (ns R)), r = 1 .. R-1, m = 0 .. ns-1, at index r ns + m - 1 (N - 1
// entries in all).  A butterfly's twiddles for one r are then consecutive across the lanes of a
// wavefront.  The angle is evaluated as 2 pi k / N with k = r m N / (ns R).
__global__ void twiddle_kernel(float2* __restrict__ tw, int N, int log2n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N - 1) return;
    const int x = e + 1;        // r ns + m
    const int n16 = log2n / 4;  // radix-16 stages
    const int st = min((31 - __clz(x)) / 4, n16);
    const int ns = 1 << (4 * st);
    const int R = (st < n16) ? 16 : (1 << (log2n % 4));
    const int r = x >> (4 * st), m = x & (ns - 1);
    const int k = r * m * (N / (ns * R));
    double s, c;
    sincos(-2.0 * kPi * k / N, &s, &c);
    tw[e] = make_float2((float)c, (float)s);
}

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Threads per FFT workgroup: 128 up to N = 2048 (one radix-16 butterfly per thread and stage,
// and twice the workgroups per CU of 256 threads), 256 above (the butterflies' registers)
constexpr int fft_threads(int log2n) { return log2n <= 11 ? 128 : 256; }

// In-register DFTs of 2, 4, 8 and 16 points (forward: exp(-2 pi i n k / R); `inv` conjugates),
// results in natural order.  8 = 2 x 4 and 16 = 4 x 4 Cooley-Tukey with the internal twiddles.
__device__ inline float2 mul_mi(float2 a, bool inv) {  // a * (-i), or a * (+i) for the inverse
    return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

__device__ inline void dft4(float2& x0, float2& x1, float2& x2, float2& x3, bool inv) {
    const float2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
    const float2 a3 = mul_mi(csub(x1, x3), inv);
    x0 = cadd(a0, a2);
    x1 = cadd(a1, a3);
    x2 = csub(a0, a2);
    x3 = csub(a1, a3);
}

template <int R>
__device__ inline void dft_reg(float2 (&v)[R], bool inv) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], inv);
    } else if constexpr (R == 8) {
        // n = 2 n1 + n2 (n1 < 4, n2 < 2), k = k1 + 4 k2
        constexpr float h = 0.70710678118654752f;
        dft4(v[0], v[2], v[4], v[6], inv);
        dft4(v[1], v[3], v[5], v[7], inv);
        // twiddles W8^(n2 k1) on the odd half: k1 = 1, 2, 3
        const float s = inv ? 1.0f : -1.0f;
        v[3] = cmul(v[3], make_float2(h, s * h));
        v[5] = mul_mi(v[5], inv);
        v[7] = cmul(v[7], make_float2(-h, s * h));
        float2 o[8];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
            o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = o[k];
    } else {
        static_assert(R == 16, "radix");
        // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, twiddle W16^(n2 k1), DFT4 over n2
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2], inv);
        const float s = inv ? 1.0f : -1.0f;
        constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
        // element [4 k1 + n2] now holds A[n2][k1]
        v[4 * 1 + 1] = cmul(v[5], make_float2(c1, s * s1));    // W16^1
        v[4 * 2 + 1] = cmul(v[9], make_float2(h, s * h));      // W16^2
        v[4 * 3 + 1] = cmul(v[13], make_float2(s1, s * c1));   // W16^3
        v[4 * 1 + 2] = cmul(v[6], make_float2(h, s * h));      // W16^2
        v[4 * 2 + 2] = mul_mi(v[10], inv);                     // W16^4
        v[4 * 3 + 2] = cmul(v[14], make_float2(-h, s * h));    // W16^6
        v[4 * 1 + 3] = cmul(v[7], make_float2(s1, s * c1));    // W16^3
        v[4 * 2 + 3] = cmul(v[11], make_float2(-h, s * h));    // W16^6
        v[4 * 3 + 3] = cmul(v[15], make_float2(-c1, -s * s1)); // W16^9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3], inv);
        // element [4 k1 + k2] holds X[k1 + 4 k2]
        float2 o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = o[k];
    }
}

// largest power of two not above r (r >= 1)
constexpr int top_bit(int r) { return (r & (r - 1)) == 0 ? r : top_bit(r & (r - 1)); }

// One radix-R Stockham stage over sub-transforms of length ns, with T threads per workgroup:
// butterfly j reads element j + r N/R, applies the twiddles exp(-+2 pi i r (j mod ns) / (ns R))
// (stage-major table of twiddle_kernel; the power-of-two ones loaded, the others formed as
// products of two earlier ones, at most three roundings deep), takes an R-point DFT and writes
// element (j / ns) ns R + j mod ns + r ns.  Elements come from the LDS buffer or, for the
// filter's first forward stage, from a global-memory loader; they go to the LDS buffer, to it
// times the spectrum (the filter's last forward stage) or, for its last inverse stage, to a
// global-memory writer.  All inputs are read into registers before a barrier, then written, so
// one LDS buffer suffices:
//   SRC 0: buf, SRC 1: load(n) (no barrier before the writes: nothing in buf is read);
//   DST 0: buf, DST 1: buf times spec[n], DST 2: store(n, value) (no barrier at all).
constexpr int kFromLds = 0, kFromGlobal = 1;
constexpr int kToLds = 0, kToLdsTimesSpectrum = 1, kToGlobal = 2;

template <int R, int N, int T, int SRC, int DST, typename Load, typename Store>
__device__ void stockham_stage_io(float2* buf, const float2* __restrict__ tw, int ns, bool inv,
                                  const double* __restrict__ spec, Load&& load, Store&& store) {
    constexpr int NBF = N / R;  // butterflies
    constexpr int NB = (NBF + T - 1) / T;
    const int tid = threadIdx.x;
    float2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * T;
        if (j < NBF) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = SRC == kFromGlobal ? load(j + r * NBF) : buf[j + r * NBF];
            const int m = j % ns;
            if (m != 0) {
                float2 w[R];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    if ((r & (r - 1)) == 0) {
                        w[r] = tw[r * ns + m - 1];
                        if (inv) w[r].y = -w[r].y;
                    } else {
                        const int hb = top_bit(r);
                        w[r] = cmul(w[hb], w[r - hb]);
                    }
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r]);
            }
            dft_reg<R>(v[b], inv);
        }
    }
    if constexpr (SRC == kFromLds && DST != kToGlobal) __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * T;
        if (j < NBF) {
            const int base = (j / ns) * ns * R + (j % ns);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int n = base + r * ns;
                if constexpr (DST == kToGlobal) {
                    store(n, v[b][r]);
                } else if constexpr (DST == kToLdsTimesSpectrum) {
                    const float h = (float)spec[n];
                    buf[n] = make_float2(v[b][r].x * h, v[b][r].y * h);
                } else {
                    buf[n] = v[b][r];
                }
            }
        }
    }
    if constexpr (DST != kToGlobal) __syncthreads();
}

// Stage st of the length-2^LOG2N transform (radix 16 for st < LOG2N / 4, then the radix of the
// remaining factor), with its source and sink.
template <int LOG2N, int ST, int SRC, int DST, typename Load, typename Store>
__device__ void fft_stage_io(float2* buf, const float2* __restrict__ tw, bool inv, const double* __restrict__ spec,
                             Load&& load, Store&& store) {
    constexpr int N = 1 << LOG2N;
    constexpr int T = fft_threads(LOG2N);
    constexpr int n16 = LOG2N / 4;
    constexpr int R = ST < n16 ? 16 : 1 << (LOG2N % 4);
    const int ns = 1 << (4 * ST);
    stockham_stage_io<R, N, T, SRC, DST>(buf, tw, ns, inv, spec, load, store);
}

// The filter's convolution: forward FFT from `load`, times the spectrum, inverse FFT to `store`,
// with the first forward stage reading global memory, the spectrum applied as the last forward
// stage writes LDS and the last inverse stage writing global memory.
template <int LOG2N, int ST = 0, typename Load, typename Store>
__device__ void fft_filter_stages(float2* buf, const float2* __restrict__ tw, const double* __restrict__ spec,
                                  Load&& load, Store&& store) {
    constexpr int NST = LOG2N / 4 + (LOG2N % 4 ? 1 : 0);  // stages per transform
    if constexpr (ST < 2 * NST) {
        constexpr bool inv = ST >= NST;
        constexpr int k = inv ? ST - NST : ST;
        constexpr int SRC = ST == 0 ? kFromGlobal : kFromLds;
        constexpr int DST = ST == NST - 1 ? kToLdsTimesSpectrum : (ST == 2 * NST - 1 ? kToGlobal : kToLds);
        fft_stage_io<LOG2N, k, SRC, DST>(buf, tw, inv, spec, load, store);
        fft_filter_stages<LOG2N, ST + 1>(buf, tw, spec, load, store);
    }
}

// Filtering() / filter_proj (python/Reconstruction_pycuda.py:223-251): two detector rows per
// complex FFT (row A real part, row B imaginary part).  The spectrum is real and even, so
// ifft(H * fft(a + i b)) = (a*g) + i (b*g) and the two rows separate without extra work.
// The cosine pre-weight R/sqrt(R^2 + kk^2 + pp^2) is evaluated at the reference's cell-edge
// coordinates kk = ((j - nu/2) du + off_u) R/SDD, pp = ((nv/2 - k) dv + off_v) R/SDD.
// ConeBeamGPU's filtering (python/ConeBeamGPU.py:296-302) weighs at cell centres (off_u =
// du R/(2 SDD), off_v = -dv R/(2 SDD)) and filters fliplr(row): the spectrum is real and
// even, so that equals the mirror image of the filtered row, written with mirror = 1.
// The FFTs are Stockham transforms in LDS, radix-16 stages (two radix-4 levels in registers per
// LDS round trip) and one radix-2, 4 or 8 stage for the remaining factor.  The first forward
// stage reads the two rows from global memory and the last inverse stage writes them back, so the
// LDS sees a write from every stage but the last and a read from every stage but the first, and
// no separate load, spectrum or store pass.
// `in` and `out` may be the same buffer (in-place filtering): a block reads its two rows before
// its first barrier and writes them after its last one, and no other block touches them.
template <int LOG2N>
__global__ void __launch_bounds__(fft_threads(LOG2N)) fdk_filter_kernel(
    const float* in, float* out, const double* __restrict__ spec,
    const float2* __restrict__ tw, int n_rows, int nu, int nv, float R, float su, float ou,
    float sv, float ov, float out_scale, int mirror) {
    constexpr int N = 1 << LOG2N;
    constexpr int T = fft_threads(LOG2N);
    __shared__ float2 buf[N];
    const int rowA = 2 * blockIdx.x;
    const int rowB = rowA + 1;
    const bool hasB = rowB < n_rows;
    const float pa = ((float)(nv / 2.0f) - (float)(rowA % nv)) * sv + ov;
    const float pb = ((float)(nv / 2.0f) - (float)(rowB % nv)) * sv + ov;
    // sample n of the weighted, zero-padded row pair (read by the first forward stage)
    auto load = [&](int n) {
        float2 val = make_float2(0.0f, 0.0f);
        if (n < nu) {
            const float kk = ((float)n - (float)nu / 2.0f) * su + ou;
            const float wa = R * rsqrtf(R * R + kk * kk + pa * pa);
            val.x = in[(size_t)rowA * nu + n] * wa;
            if (hasB) {
                const float wb = R * rsqrtf(R * R + kk * kk + pb * pb);
                val.y = in[(size_t)rowB * nu + n] * wb;
            }
        }
        return val;
    };
    // sample n of the filtered pair (written by the last inverse stage; the padding is dropped)
    auto store = [&](int n, float2 val) {
        if (n < nu) {
            const int j = mirror ? nu - 1 - n : n;
            out[(size_t)rowA * nu + j] = val.x * out_scale;
            if (hasB) out[(size_t)rowB * nu + j] = val.y * out_scale;
        }
    };
    fft_filter_stages<LOG2N>(buf, tw, spec, load, store);
}

template <int L>
hipError_t launch_filter(int l2, dim3 grid, hipStream_t st, const float* in, float* out,
                         const double* spec, const float2* tw, int n_rows, int nu, int nv, float R,
                         float su, float ou, float sv, float ov, float scale, int mirror) {
    if (l2 == L) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(fdk_filter_kernel<L>), grid, dim3(fft_threads(L)), 0, st,
                           in, out, spec, tw, n_rows, nu, nv, R, su, ou, sv, ov, scale, mirror);
        return hipGetLastError();
    }
    if constexpr (L < kMaxLog2N) {
        return launch_filter<L + 1>(l2, grid, st, in, out, spec, tw, n_rows, nu, nv, R, su, ou, sv,
                                    ov, scale, mirror);
    }
    return hipErrorInvalidValue;
}

// Scratch memory of ctr_dd_backproject_static, for callers that cannot allocate: the FDK
// backprojection's per-view table and its (tile, view) flags.  The HIP runtime gives each GPU
// its own copy when it loads the code object there, and keeps it until the process ends.  In
// host code the name `scratch` is only a handle: its device address on the current GPU comes
// from hipGetSymbolAddress, looked up once per device and kept in scratch_ptr.
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void* scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (looked up on the first call there).
int static_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return CTR_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return CTR_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CTR_ERR_HIP;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return CTR_OK;
}

}  // namespace

// ==========================================================================================
// C ABI
// ==========================================================================================

extern "C" {

int ctr_version(void) { return CTR_ABI_VERSION; }

const char* ctr_strerror(int code) {
    switch (code) {
        case CTR_OK: return "ok";
        case CTR_ERR_INVALID: return "invalid argument (geometry, pointer or size out of range)";
        case CTR_ERR_UNSUPPORTED: return "unsupported configuration (z-slab branch, detector shape or FFT length)";
        case CTR_ERR_WORKSPACE: return "workspace too small";
        case CTR_ERR_HIP: return "HIP runtime error";
        case CTR_ERR_SCRATCH:
            return "problem too large for the static scratch of ctr_dd_backproject_static (32 MiB: "
                   "64 bytes per view plus one byte per 16x16x16 voxel tile and view)";
        default: return "unknown error";
    }
}

int ctr_select_branches(const ctr_geometry* g, const double* angles, int32_t* branch_out) {
    int st = check_geometry(g);
    if (st) return st;
    if (!angles || !branch_out) return CTR_ERR_INVALID;
    for (int i = 0; i < g->n_views; ++i) {
        const double th = angles[i];
        const double z = g->helical * th / (2.0 * kPi);
        const double S[3] = {-g->sad * std::sin(th), g->sad * std::cos(th), g->source_init[2] + z};
        const double D[3] = {(g->sdd - g->sad) * std::sin(th), -(g->sdd - g->sad) * std::cos(th),
                             g->detector_init[2] + z};
        const double ax = std::fabs(S[0] - D[0]), ay = std::fabs(S[1] - D[1]),
                     az = std::fabs(S[2] - D[2]);
        if (ax >= ay && ax >= az) {
            branch_out[i] = CTR_BRANCH_X;
        } else if (ay >= ax && ay >= az) {
            branch_out[i] = CTR_BRANCH_Y;
        } else {
            branch_out[i] = CTR_BRANCH_Z;
        }
    }
    return CTR_OK;
}

int ctr_zero_padded_length(int nu) {
    if (nu < 2) return -1;
    const int target = 2 * (nu - 1);
    int n = 1;
    while (n < target) n <<= 1;
    return n;
}

size_t ctr_dd_forward_workspace_bytes(const ctr_geometry* g) {
    if (check_geometry(g)) return 0;
    return fp_layout(g).total;
}

int ctr_dd_forward(const ctr_geometry* g, const double* angles, const int32_t* branch,
                   const float* vol, float* proj, void* workspace, size_t workspace_bytes,
                   void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (!angles || !branch || !vol || !proj || !workspace) return CTR_ERR_INVALID;
    const FpLayout L = fp_layout(g);
    if (workspace_bytes < L.total) return CTR_ERR_WORKSPACE;
    // the projector addresses a batch of 64 slabs of the z-prefix tables with 31-bit offsets
    if (64.0 * 4.0 * zprefix_stride(g->nz) * std::max(g->nx, g->ny) >= 2147483647.0)
        return CTR_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    float* tabs = (float*)(ws + L.tabs_off);
    float* rows = (float*)(ws + L.rows_off);
    float* t_ys = (float*)(ws + L.volt_off);
    float* t_xs = (float*)(ws + L.volt_off + L.copy_bytes);
    const int stride = fp_view_stride(g->nu);
    const int gx = (std::max(g->nu + 1, g->nv + 1) + 255) / 256;
    hipLaunchKernelGGL(fp_geometry_kernel, dim3(gx, g->n_views), dim3(256), 0, s, *g, angles, branch,
                       tabs, rows, stride);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    const int nzp = zprefix_stride(g->nz);
    hipLaunchKernelGGL(zprefix_kernel, dim3((g->nx + 31) / 32, g->ny), dim3(256), 0, s, vol, t_ys, t_xs,
                       g->nx, g->ny, g->nz, nzp);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    const int col_groups = (g->nu + kFpCols * kFpWaves - 1) / (kFpCols * kFpWaves);
    const int row_blocks = (g->nv + kColRows - 1) / kColRows;
    const int n_pairs = (g->n_views + kFpViews - 1) / kFpViews;
    const int n_vblocks = (n_pairs + kViewBlock - 1) / kViewBlock;
    const long long n_blocks =
        8LL * ((col_groups + 7) / 8) * row_blocks * n_vblocks * kViewBlock;
    if (n_blocks >= 2147483647LL) return CTR_ERR_INVALID;
    hipLaunchKernelGGL(dd_forward_kernel, dim3((unsigned)n_blocks), dim3(64 * kFpWaves), 0, s, t_ys, t_xs,
                       proj, tabs, rows, branch, g->nx, g->ny, g->nz, nzp, g->nu, g->nv, stride, 1e-6f,
                       g->n_views, col_groups, n_vblocks);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    // z-slab views (left at zero above): their own kernel, which returns at once for the others
    hipLaunchKernelGGL(dd_forward_zslab_kernel, dim3(kZslabGroups, g->n_views), dim3(256), 0, s, *g, angles,
                       branch, vol, proj, 1e-6);
    return status_from_hip(hipGetLastError());
}

size_t ctr_fdk_filter_workspace_bytes(const ctr_geometry* g) {
    if (check_geometry(g)) return 0;
    const int N = ctr_zero_padded_length(g->nu);
    if (N <= 0) return 0;
    return align_up((size_t)N * sizeof(float2), 256);
}

int ctr_filter_spectrum(const ctr_geometry* g, int filter_type, double cutoff, double* spectrum,
                        void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (!spectrum) return CTR_ERR_INVALID;
    if (filter_type < CTR_FILTER_RAM_LAK || filter_type > CTR_FILTER_HANN) return CTR_ERR_INVALID;
    if (filter_type != CTR_FILTER_RAM_LAK && !(cutoff > 0)) return CTR_ERR_INVALID;
    const int N = ctr_zero_padded_length(g->nu);
    if (N <= 0) return CTR_ERR_INVALID;
    if (N > (1 << kMaxLog2N)) return CTR_ERR_UNSUPPORTED;
    const double tau = g->du * g->sad / g->sdd;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(spectrum_kernel, dim3((N + 255) / 256), dim3(256), N * sizeof(double), s,
                       spectrum, N, tau, filter_type, cutoff, 1.0, kPi * cutoff / (2.0 * tau));
    return status_from_hip(hipGetLastError());
}

int ctr_fdk_filter(const ctr_geometry* g, const double* spectrum, const float* proj_in,
                   float* proj_out, int n_views, void* workspace, size_t workspace_bytes,
                   void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (!spectrum || !proj_in || !proj_out || !workspace || n_views <= 0) return CTR_ERR_INVALID;
    if ((double)n_views * g->nv * g->nu >= 2147483647.0) return CTR_ERR_INVALID;
    const int N = ctr_zero_padded_length(g->nu);
    if (N <= 0) return CTR_ERR_INVALID;
    const int l2 = log2_exact(N);
    if (l2 < 3) return CTR_ERR_UNSUPPORTED;  // nu >= 5
    if (l2 > kMaxLog2N) return CTR_ERR_UNSUPPORTED;
    if (workspace_bytes < ctr_fdk_filter_workspace_bytes(g)) return CTR_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    float2* tw = (float2*)workspace;
    hipLaunchKernelGGL(twiddle_kernel, dim3((N + 255) / 256), dim3(256), 0, s, tw, N, l2);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    const int n_rows = n_views * g->nv;
    const double R = g->sad;
    const double mag = g->sad / g->sdd;  // R / (R + D)
    const double tau = g->du * mag;
    const dim3 grid((n_rows + 1) / 2);
    return status_from_hip(launch_filter<3>(
        l2, grid, s, proj_in, proj_out, spectrum, tw, n_rows, g->nu, g->nv, (float)R,
        (float)(g->du * mag), (float)(g->det_offset[0] * mag), (float)(g->dv * mag),
        (float)(g->det_offset[1] * mag), (float)(tau / N), 0));
}

int ctr_pixel_filter_spectrum(int n_fft, double tau, int filter_type, double cutoff,
                              double* spectrum, void* stream) {
    if (!spectrum || !(tau > 0)) return CTR_ERR_INVALID;
    if (filter_type < CTR_FILTER_RAM_LAK || filter_type > CTR_FILTER_HANN) return CTR_ERR_INVALID;
    if (filter_type != CTR_FILTER_RAM_LAK && !(cutoff > 0)) return CTR_ERR_INVALID;
    if (!(cutoff >= 0)) return CTR_ERR_INVALID;
    const int l2 = log2_exact(n_fft);
    if (l2 < 1) return CTR_ERR_INVALID;
    if (l2 > kMaxLog2N) return CTR_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(spectrum_kernel, dim3((n_fft + 255) / 256), dim3(256),
                       n_fft * sizeof(double), s, spectrum, n_fft, tau, filter_type, cutoff, 2.0,
                       kPi * cutoff);
    return status_from_hip(hipGetLastError());
}

int ctr_pixel_filter(const ctr_geometry* g, const double* spectrum, const float* proj_in,
                     float* q_out, int n_views, void* workspace, size_t workspace_bytes,
                     void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (g->detector_shape != CTR_DETECTOR_FLAT) return CTR_ERR_UNSUPPORTED;
    if (!spectrum || !proj_in || !q_out || !workspace || n_views <= 0) return CTR_ERR_INVALID;
    if (proj_in == q_out) return CTR_ERR_INVALID;  // rows are written mirrored
    if ((double)n_views * g->nv * g->nu >= 2147483647.0) return CTR_ERR_INVALID;
    const int N = ctr_zero_padded_length(g->nu);
    if (N <= 0) return CTR_ERR_INVALID;
    const int l2 = log2_exact(N);
    if (l2 < 3) return CTR_ERR_UNSUPPORTED;  // nu >= 5
    if (l2 > kMaxLog2N) return CTR_ERR_UNSUPPORTED;
    if (workspace_bytes < ctr_fdk_filter_workspace_bytes(g)) return CTR_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    float2* tw = (float2*)workspace;
    hipLaunchKernelGGL(twiddle_kernel, dim3((N + 255) / 256), dim3(256), 0, s, tw, N, l2);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    const int n_rows = n_views * g->nv;
    const double mag = g->sad / g->sdd;
    const double su = g->du * mag, sv = g->dv * mag;
    const dim3 grid((n_rows + 1) / 2);
    return status_from_hip(launch_filter<3>(l2, grid, s, proj_in, q_out, spectrum, tw, n_rows,
                                            g->nu, g->nv, (float)g->sad, (float)su,
                                            (float)(0.5 * su), (float)sv, (float)(-0.5 * sv),
                                            (float)(1.0 / N), 1));
}

// Adjoint workspace: the forward projector's tables and row table (no volume tables), the
// per-view adjoint constants, then the row-prefix table Y[view][column][nvp].
struct AdjLayout {
    size_t tabs_off, rows_off, at_off, yl_off, total;
};

// entries per column of the adjoint's row-prefix table: breakpoints e = 0 .. nv and a copy
// of the total, so that the pair (k, k+1) around any point of [0, nv] is inside the column
int adj_prefix_stride(int nv) { return nv + 2; }

AdjLayout adj_layout(const ctr_geometry* g) {
    AdjLayout L;
    L.tabs_off = 0;
    L.rows_off = align_up((size_t)fp_view_stride(g->nu) * g->n_views * 4, 256);
    L.at_off = align_up(L.rows_off + (size_t)(2 * g->nv + 1) * 4, 256);
    L.yl_off = align_up(L.at_off + (size_t)g->n_views * kAdjViewStride * 4, 256);
    L.total = align_up(L.yl_off + (size_t)g->n_views * g->nu * adj_prefix_stride(g->nv) * 4, 256);
    return L;
}

// FDK backprojection workspace: per-view constants, then one flag byte per (tile, view)
size_t bp_tiles(const ctr_geometry* g) {
    return (size_t)((g->nx + kTileX - 1) / kTileX) * ((g->ny + kTileY - 1) / kTileY) *
           ((g->nz + kBpZC - 1) / kBpZC);
}

size_t ctr_dd_backproject_workspace_bytes(const ctr_geometry* g, int mode) {
    if (check_geometry(g)) return 0;
    if (mode == CTR_BP_ADJOINT) return adj_layout(g).total;
    return align_up((size_t)g->n_views * kBpViewStride * sizeof(float), 256) +
           align_up(bp_tiles(g) * g->n_views, 256);
}

int ctr_dd_backproject(const ctr_geometry* g, const double* angles, const int32_t* branch,
                       const float* proj, float* vol, int mode, float scale, float zero_tol,
                       int accumulate, void* workspace, size_t workspace_bytes, void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (!angles || !proj || !vol || !workspace) return CTR_ERR_INVALID;
    if (mode != CTR_BP_FDK_REF && mode != CTR_BP_FDK_EDGE && mode != CTR_BP_ADJOINT)
        return CTR_ERR_INVALID;
    if (workspace_bytes < ctr_dd_backproject_workspace_bytes(g, mode)) return CTR_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    if (mode == CTR_BP_ADJOINT) {
        // transpose of ctr_dd_forward: its weights (and its 1e-6 transverse weight zeroing) by construction
        if (!branch) return CTR_ERR_INVALID;
        const AdjLayout L = adj_layout(g);
        char* ws = (char*)workspace;
        float* tabs = (float*)(ws + L.tabs_off);
        float* rows = (float*)(ws + L.rows_off);
        float* at = (float*)(ws + L.at_off);
        float* yp = (float*)(ws + L.yl_off);  // row-prefix table
        const int stride = fp_view_stride(g->nu);
        const int gx = (std::max(g->nu + 1, g->nv + 1) + 255) / 256;
        hipLaunchKernelGGL(fp_geometry_kernel, dim3(gx, g->n_views), dim3(256), 0, s, *g, angles,
                           branch, tabs, rows, stride);
        if ((st = status_from_hip(hipGetLastError()))) return st;
        hipLaunchKernelGGL(adj_geometry_kernel, dim3((g->n_views + 255) / 256), dim3(256), 0, s, *g,
                           angles, branch, at);
        if ((st = status_from_hip(hipGetLastError()))) return st;
        const int nvp = adj_prefix_stride(g->nv);
        hipLaunchKernelGGL(adj_rowprefix_kernel, dim3((g->nu + 31) / 32, g->n_views), dim3(256), 0, s, proj,
                           yp, tabs, rows, g->nu, g->nv, nvp, stride);
        if ((st = status_from_hip(hipGetLastError()))) return st;
        const dim3 block(64, 4);
        const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, (g->nz + kAdjZChunk - 1) / kAdjZChunk);
        hipLaunchKernelGGL(dd_adjoint_kernel, grid, block, 0, s, yp, vol, tabs, at, g->nx, g->ny, g->nz,
                           g->nu, g->nv, nvp, g->n_views, stride,
                           g->detector_shape == CTR_DETECTOR_CURVED ? 1 : 0, scale, 1e-6f, accumulate);
        return status_from_hip(hipGetLastError());
    }
    float* vt = (float*)workspace;
    hipLaunchKernelGGL(bp_geometry_kernel, dim3((g->n_views + 255) / 256), dim3(256), 0, s, *g,
                       angles, vt, mode);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    // voxel centres: ReconCenter + (i - (n-1)/2) * d with d = (dx, -dy, -dz)
    const float x0 = (float)(g->recon_center[0] - (g->nx - 1) / 2.0 * g->dx);
    const float y0 = (float)(g->recon_center[1] + (g->ny - 1) / 2.0 * g->dy);
    const float z0 = (float)(g->recon_center[2] + (g->nz - 1) / 2.0 * g->dz);
    constexpr int ZC = kBpZC;
    const int tiles_x = (g->nx + kTileX - 1) / kTileX;
    const int tiles_y = (g->ny + kTileY - 1) / kTileY;
    const int tiles_z = (g->nz + ZC - 1) / ZC;
    const long long tiles = (long long)tiles_x * tiles_y * tiles_z;
    if (tiles >= 2147483647LL) return CTR_ERR_INVALID;
    const int curved = g->detector_shape == CTR_DETECTOR_CURVED ? 1 : 0;
    const float fdx = (float)g->dx, fdy = (float)(-g->dy), fdz = (float)(-g->dz);
    // weight zeroing, a volume of one z chunk (a slice through the source height, as in a 2D
    // fan, is on neither side of it), or a caller's request for one pass (zero_tol < 0): the
    // general kernel does every view
    if (zero_tol != 0.0f || tiles_z == 1) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(dd_backproject_tile_kernel<ZC, kBpAll>), dim3((unsigned)tiles),
                           dim3(256), 0, s, proj, vol, vt, (uint8_t*)nullptr, g->nx, g->ny, g->nz, g->nu,
                           g->nv, g->n_views, x0, y0, z0, fdx, fdy, fdz, curved, scale, zero_tol, accumulate,
                           tiles_x, tiles_y, tiles_z);
        return status_from_hip(hipGetLastError());
    }
    // fast pass over the common case, then the (tile, view) pairs it left
    uint8_t* flags = (uint8_t*)workspace + align_up((size_t)g->n_views * kBpViewStride * sizeof(float), 256);
    const size_t flag_words = align_up((size_t)tiles * g->n_views, 256) / 4;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<size_t>((flag_words + 255) / 256, 65536)),
                       dim3(256), 0, s, (uint32_t*)flags, flag_words);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(dd_backproject_tile_kernel<ZC, kBpFast>), dim3((unsigned)tiles),
                       dim3(256), 0, s, proj, vol, vt, flags, g->nx, g->ny, g->nz, g->nu, g->nv, g->n_views,
                       x0, y0, z0, fdx, fdy, fdz, curved, scale, zero_tol, accumulate, tiles_x, tiles_y, tiles_z);
    if ((st = status_from_hip(hipGetLastError()))) return st;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(dd_backproject_tile_kernel<ZC, kBpRest>), dim3((unsigned)tiles),
                       dim3(256), 0, s, proj, vol, vt, flags, g->nx, g->ny, g->nz, g->nu, g->nv, g->n_views,
                       x0, y0, z0, fdx, fdy, fdz, curved, scale, zero_tol, accumulate, tiles_x, tiles_y, tiles_z);
    return status_from_hip(hipGetLastError());
}

int ctr_dd_backproject_static(const ctr_geometry* g, const double* angles, const float* proj,
                              float* vol, int mode, float scale, float zero_tol, int accumulate,
                              void* stream) {
    int st = check_geometry(g);
    if (st) return st;
    if (mode != CTR_BP_FDK_REF && mode != CTR_BP_FDK_EDGE) return CTR_ERR_UNSUPPORTED;
    const size_t need = ctr_dd_backproject_workspace_bytes(g, mode);
    if (need > kStaticScratchBytes) return CTR_ERR_SCRATCH;
    void* ws = nullptr;
    if ((st = static_scratch(&ws))) return st;
    // the same launches as ctr_dd_backproject: the per-view table is written by a kernel and
    // the flags are cleared by a fill kernel on `stream` before use
    return ctr_dd_backproject(g, angles, nullptr, proj, vol, mode, scale, zero_tol, accumulate, ws,
                              need, stream);
}

}  // extern "C"
