/*
 * ctrecon.h -- C ABI of the MI355X (gfx950) distance-driven CT projector pair and FDK filter.
 *
 * Every buffer is caller-owned device memory (C-contiguous float32 / float64 / int32);
 * the library never allocates: scratch memory comes in through a workspace argument whose
 * size the *_workspace_bytes() queries return.  Work is enqueued on the HIP stream passed
 * as `stream` (a hipStream_t, NULL = default stream); calls return before the GPU finishes.
 * Every function returns CTR_OK (0) or a CTR_ERR_* code; ctr_strerror() names it.  Invalid
 * arguments are rejected on the host before anything is launched.
 *
 * Array layouts are the reference's (leehoy/CTReconstruction, python/Reconstruction_pycuda.py):
 *   volume       float32 [nz][ny][nx]            (LoadRecon, :104-105)
 *   projections  float32 [n_views][nv][nu]       (LoadProj,  :101-102)
 *   angles       float64 [n_views]               (linspace(StartAngle, EndAngle, n+1)[:-1], :788-789)
 *   branch       int32   [n_views]               (slab axis per view, :852-853 / :921-922)
 */
#ifndef CTRECON_H
#define CTRECON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTR_ABI_VERSION 1

enum ctr_status {
    CTR_OK = 0,
    CTR_ERR_INVALID = 1,      /* bad geometry, null pointer, size out of range            */
    CTR_ERR_UNSUPPORTED = 2,  /* unknown detector shape, FFT length > 8192, sizes past 31-bit offsets */
    CTR_ERR_WORKSPACE = 3,    /* workspace smaller than the *_workspace_bytes() query     */
    CTR_ERR_HIP = 4,          /* a HIP launch or runtime call failed                       */
    CTR_ERR_SCRATCH = 5       /* too large for ctr_dd_backproject_static's static scratch  */
};

enum ctr_detector_shape { CTR_DETECTOR_FLAT = 0, CTR_DETECTOR_CURVED = 1 };

enum ctr_branch { CTR_BRANCH_X = 0, CTR_BRANCH_Y = 1, CTR_BRANCH_Z = 2 };

enum ctr_filter_type {
    CTR_FILTER_RAM_LAK = 0, CTR_FILTER_SHEPP_LOGAN = 1, CTR_FILTER_COSINE = 2,
    CTR_FILTER_HAMMING = 3, CTR_FILTER_HANN = 4
};

enum ctr_bp_mode {
    CTR_BP_FDK_REF = 0,   /* reference grid, half a pixel off the forward edges (:353-354)   */
    CTR_BP_FDK_EDGE = 1,  /* forward projector's true cell edges (:356-357, commented there) */
    CTR_BP_ADJOINT = 2    /* exact transpose of ctr_dd_forward (no reference counterpart)    */
};

/* Geometry of one scan, in the reference's params-dict conventions
 * (python/Reconstruction_pycuda.py:64-99).  Spacings are positive; the library applies the
 * reference's sign flips (dy, dz, dv -> negative) itself. */
typedef struct ctr_geometry {
    int32_t nx, ny, nz;        /* NumberOfImage                                     */
    int32_t nu, nv;            /* NumberOfDetectorPixels                            */
    int32_t n_views;           /* NumberOfViews                                     */
    int32_t detector_shape;    /* ctr_detector_shape (DetectorShape)                */
    int32_t reserved0;
    double dx, dy, dz;         /* ImagePixelSpacing [mm]                            */
    double du, dv;             /* DetectorPixelSize [mm] (curved: arc length at SDD) */
    double sad;                /* |SourceInit - RotationOrigin|                     */
    double sdd;                /* |SourceInit - DetectorInit|                       */
    double source_init[3];     /* SourceInit (the backprojector uses it unrotated)  */
    double detector_init[3];   /* DetectorInit                                      */
    double helical;            /* Pitch * nv * dv * SAD / SDD: z advance per 2*pi   */
    double recon_center[3];    /* ReconCenter                                       */
    double det_offset[2];      /* DetectorOffset (u, v)                             */
} ctr_geometry;

/* Library ABI version (CTR_ABI_VERSION of the build). */
int ctr_version(void);

/* Static message for a ctr_status code.  Replaces the reference's ErrorDescription
 * (python/Reconstruction_pycuda.py:43-59), which printed and called sys.exit(). */
const char* ctr_strerror(int code);

/* Host helper: slab branch of every view from the reference's FP64 comparisons
 * (python/Reconstruction_pycuda.py:828-830, 852-853, 921-922; ties go to x, then y, then z).
 * `angles` and `branch_out` are HOST arrays of n_views entries. */
int ctr_select_branches(const ctr_geometry* g, const double* angles, int32_t* branch_out);

/* Zero-padded FFT length 2^ceil(log2(2(nu-1))) (python/Reconstruction_pycuda.py:236). */
int ctr_zero_padded_length(int nu);

/* ---- DD forward projection --------------------------------------------------------------
 * Replaces Reconstruction.distance_forward's per-view host geometry and H2D copies
 * (python/Reconstruction_pycuda.py:762-1059) and the kernels distance_project_on_{x,y,z}3
 * (python/GPUFuncs_pycuda.py:398-587).  Per-view footprint tables are built on the GPU in
 * FP64; the projector is a gather (slabs summed in registers per detector cell, no atomics,
 * deterministic).  z-slab views (the z branch, which only non-circular source/detector offsets
 * select) go through a plain FP64 kernel whose cost grows with the footprint area.  `proj` is
 * overwritten.  `vol` and `proj` must not
 * overlap the workspace.  CTR_ERR_UNSUPPORTED when 256 * (nz + 2) * max(nx, ny) >= 2^31
 * (e.g. above 2048 x 2048 x 2046). */
size_t ctr_dd_forward_workspace_bytes(const ctr_geometry* g);
int ctr_dd_forward(const ctr_geometry* g, const double* angles, const int32_t* branch,
                   const float* vol, float* proj, void* workspace, size_t workspace_bytes,
                   void* stream);

/* ---- FDK filter -------------------------------------------------------------------------
 * ctr_filter_spectrum replaces Reconstruction.Filter (python/Reconstruction_pycuda.py:186-221):
 * writes the filter magnitude for an FFT of length N = ctr_zero_padded_length(nu) in FFT
 * (ifftshift-ed) order, float64[N], computed in FP64 on the GPU from the cosine series of
 * the discrete Ram-Lak taps, with the reference's windows, its Shepp-Logan DC zero and its
 * frequency mask |w| > pi*cutoff/(2*tau), tau = du*SAD/SDD.
 *
 * ctr_fdk_filter replaces Filtering()/filter_proj (python/Reconstruction_pycuda.py:223-251):
 * cosine pre-weight at cell-edge coordinates, zero-padded FFT convolution with `spectrum`,
 * crop to nu, times tau.  `proj_in` may equal `proj_out` (in place). */
size_t ctr_fdk_filter_workspace_bytes(const ctr_geometry* g);
int ctr_filter_spectrum(const ctr_geometry* g, int filter_type, double cutoff,
                        double* spectrum, void* stream);
int ctr_fdk_filter(const ctr_geometry* g, const double* spectrum, const float* proj_in,
                   float* proj_out, int n_views, void* workspace, size_t workspace_bytes,
                   void* stream);

/* ---- DD backprojection ------------------------------------------------------------------
 * Replaces Reconstruction.distance_backproj (python/Reconstruction_pycuda.py:320-435) and
 * the kernel flat_distance_backproj_arb2 (python/GPUFuncs_pycuda.py:837-956), or for
 * curved detectors curved_distance_backproj_arb with the numba units
 * (python/GPUFuncs_numba.py:537-685, python/Reconstruction_numba.py:399-413).
 * Voxel-driven gather over all views in one launch, no atomics.
 * vol = (accumulate ? vol : 0) + sum_views scale * BP_view(proj); the reference's FDK uses
 * scale = dtheta (Q = proj * dtheta, :411).  Overlap weights below `zero_tol` are zeroed
 * (0 = never, as the pycuda kernel; 1e-5 = the numba kernels).  A negative `zero_tol` zeroes
 * nothing either, but runs every (voxel, view) pair in one pass instead of a common-case pass
 * followed by a completing pass (the same sums in another order: a check of that split).
 * mode CTR_BP_ADJOINT computes scale * A^T proj, the exact transpose of ctr_dd_forward (its
 * weights and its 1e-6 transverse zeroing; `zero_tol` is ignored) over the x- and y-slab views;
 * z-slab views add nothing.  It needs `branch` and has no counterpart in the reference, which
 * only has the FDK-weighted backprojector. */
size_t ctr_dd_backproject_workspace_bytes(const ctr_geometry* g, int mode);
int ctr_dd_backproject(const ctr_geometry* g, const double* angles, const int32_t* branch,
                       const float* proj, float* vol, int mode, float scale, float zero_tol,
                       int accumulate, void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference: ctr_dd_backproject (modes CTR_BP_FDK_REF and
 * CTR_BP_FDK_EDGE) without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch (the per-view table and the per-(tile, view) flags) lives in a static 32 MiB
 * device array, one per GPU, held until the process ends; problems needing more
 * (ctr_dd_backproject_workspace_bytes(g, mode) > 32 MiB, e.g. above about 1000 views at 512^3)
 * return CTR_ERR_SCRATCH, and CTR_BP_ADJOINT returns CTR_ERR_UNSUPPORTED.  It allocates
 * nothing and launches the same kernels as ctr_dd_backproject.  Every call on one GPU shares
 * that GPU's array: calls that overlap in time on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  The current device must be the one `stream` and the
 * buffers belong to. */
int ctr_dd_backproject_static(const ctr_geometry* g, const double* angles, const float* proj,
                              float* vol, int mode, float scale, float zero_tol, int accumulate,
                              void* stream);

/* ---- Pixel-driven cone-beam FDK ----------------------------------------------------------
 * Replaces ConeBeam.Reconstruction of python/ConeBeamGPU.py:203-321, the reference's
 * voxel-driven FDK with bilinear detector interpolation.  Geometry fields used: nx, ny, nz,
 * nu, nv, n_views, dx, dy, dz, du, dv, sad (R), sdd, recon_center; detector_shape must be
 * flat and det_offset, helical zero (CTR_ERR_UNSUPPORTED otherwise); nu, nv >= 2.  Voxel
 * (ix, iy, iz) sits at ((ix - (nx-1)/2) dx, (iy - (ny-1)/2) dy, (iz - (nz-1)/2) dz) +
 * recon_center, every axis ascending: the reference's linspace(-fov/2, fov/2, n) grid when
 * dx = fov/(n-1).  Detector cell (k, j) of a view sits at ((j - (nu-1)/2) du, (k - (nv-1)/2) dv)
 * scaled by R/SDD to the rotation axis.
 *
 * ctr_pixel_filter_spectrum replaces ConeBeam.Filter (:355-384): 2 |DFT| of the Ram-Lak taps
 * for pixel size tau, windowed, zero where |w| > pi*cutoff (w in rad/sample), in FFT order,
 * float64[n_fft]; n_fft a power of two <= 8192.
 *
 * ctr_pixel_filter replaces the per-view filtering (:243, :296-302): weight
 * R/sqrt(R^2 + kk^2 + pp^2) at cell centres, fliplr, zero-padded FFT convolution with
 * `spectrum` (length ctr_zero_padded_length(nu)), crop to nu, no tau factor.  The output Q is
 * the reference's: its columns are mirrored (Q[v][k][j] comes from input column nu-1-j).
 * Workspace: ctr_fdk_filter_workspace_bytes(g).  proj_in must differ from q_out.
 *
 * ctr_pixel_backproject replaces CalculateInterpolGrid + Interpol2dLineargpu and the view loop
 * (:65-129, :303-315): vol = (accumulate ? vol : 0) + sum_views scale * R^2/(R-s)^2 *
 * bilinear(Q_view; R t/(R-s), R z/(R-s)) with t = x cos + y sin, s = -x sin + y cos and the
 * sample 0 outside the detector grid (the reference's fill_value); the reference's scale is
 * dtheta.  The source must stay outside the voxel grid (R - s > 0 everywhere). */
int ctr_pixel_filter_spectrum(int n_fft, double tau, int filter_type, double cutoff,
                              double* spectrum, void* stream);
int ctr_pixel_filter(const ctr_geometry* g, const double* spectrum, const float* proj_in,
                     float* q_out, int n_views, void* workspace, size_t workspace_bytes,
                     void* stream);
size_t ctr_pixel_backproject_workspace_bytes(const ctr_geometry* g);
int ctr_pixel_backproject(const ctr_geometry* g, const double* angles, const float* q,
                          float* vol, float scale, int accumulate, void* workspace,
                          size_t workspace_bytes, void* stream);

/* ---- Learned prior: 3x3 convolutions, bfloat16 on MFMA, FP32 accumulation ---------------
 * No counterpart in the reference (the reference has no learned prior).
 * Activations are NHWC bfloat16 [n][h][w][c]; weights bfloat16 [cout][3][3][cin]; bias float32.
 * ctr_conv3x3_bf16_forward: out = act(conv(in, weight) + bias) + residual ("same" zero
 * padding; act = ReLU if relu != 0; bias and residual may be NULL); cout <= 64.
 * The input gradient is ctr_conv3x3_bf16_forward applied to the output gradient with the
 * weights flipped in space and transposed in channels.
 * ctr_conv3x3_bf16_wgrad: grad_weight[cout][3][3][cin] (float32) and grad_bias[cout] (may be
 * NULL) summed over all pixels, in a fixed order; cin, cout <= 64. */
int ctr_conv3x3_bf16_forward(const void* in, const void* weight, const float* bias,
                             const void* residual, void* out, int n, int h, int w, int cin,
                             int cout, int relu, void* stream);
size_t ctr_conv3x3_bf16_wgrad_workspace_bytes(int n, int h, int w, int cin, int cout);
int ctr_conv3x3_bf16_wgrad(const void* in, const void* grad_out, float* grad_weight,
                           float* grad_bias, int n, int h, int w, int cin, int cout,
                           void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CTRECON_H */
