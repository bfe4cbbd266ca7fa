/*
 * zodi_hip.h — C ABI of the MI355X (gfx950) zodiacal line-of-sight integral.
 *
 * Plain C types only: device pointers, sizes and an opaque HIP stream handle.
 * Every entry point returns a zodi_status (0 on success). All buffers are owned
 * by the caller and live in device memory unless stated otherwise; calls are
 * asynchronous on `stream` (NULL = the default stream).
 *
 * Reference interfaces replaced (paths relative to the zodipy 1.1.3 tree):
 *   zodi_forward          zodipy/model.py:253-277   (get_line_of_sight_range +
 *                         the components x Gauss-Legendre nodes loop of
 *                         Model._evaluate, i.e. zodipy/line_of_sight.py:55-105,
 *                         zodipy/brightness.py:21-83, zodipy/number_density.py:47-404,
 *                         zodipy/scattering.py:11-59, zodipy/blackbody.py:16-30)
 *   zodi_line_of_sight    zodipy/line_of_sight.py:88-105 (per-component start/stop)
 *   zodi_number_density   zodipy/number_density.py:47-420 (component densities at
 *                         heliocentric points; the kernel behind grid_number_density,
 *                         zodipy/number_density.py:482-536)
 *   zodi_pix2vec,         zodipy/model.py:247-251 (the frame rotation and cartesian
 *   zodi_lonlat_to_vec    conversion of the pointings; HEALPix pixel centres)
 *   zodi_spline_positions zodipy/bodies.py:22-35 (time-ordered body positions)
 *   zodi_max_radius       zodipy/line_of_sight.py:71-73 (the any-observer cutoff test)
 *   zodi_adjoint          no reference counterpart: vector-Jacobian product of the
 *                         forward map with respect to the model parameter vector
 *                         (the order of zodipy/zodiacal_light_model.py:38-55's
 *                         parameter dictionary, followed by the effective spectral
 *                         parameters), for parameter fitting
 *   zodi_jacobian         no reference counterpart: the per-ray Jacobian of the
 *                         same map, for Gauss-Newton and Fisher-matrix work
 *   zodi_table_vjp        no reference counterpart: the gradient with respect to
 *                         the Planck table (zodipy/blackbody.py:33-49), for fits of
 *                         the bandpass weights
 *   zodi_spectral_vjp     no reference counterpart: the same together with the
 *                         gradient with respect to the effective spectral
 *                         parameters, for bandpass and spectral-table fits
 *   zodi_spectral_setup   zodipy/unpack_model.py:23-173 and zodipy/blackbody.py:33-49
 *                         (effective spectral parameters and the Planck table,
 *                         written into a device model block)
 */
#ifndef ZODI_HIP_H
#define ZODI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZODI_ABI_VERSION 2

#define ZODI_MAX_COMPS 16
#define ZODI_MAX_GROUPS 16
#define ZODI_MAX_NODES 512
#define ZODI_N_TEMPS 100
#define ZODI_COMP_SLOTS 32
#define ZODI_MAX_PARAMS 256
#define ZODI_SHARED_SLOTS 8

typedef int32_t zodi_status;
enum {
  ZODI_OK = 0,
  ZODI_ERR_ARG = 1,        /* invalid argument (null pointer, size, mode) */
  ZODI_ERR_MODEL = 2,      /* malformed model block (version, counts) */
  ZODI_ERR_HIP = 3,        /* a HIP runtime call failed; see zodi_last_error() */
  ZODI_ERR_TOO_LARGE = 4   /* size beyond a documented limit */
};

/* Integrand families (zodipy/brightness.py:21-56 and :59-83). */
enum { ZODI_INTEGRAND_KELSALL = 0, ZODI_INTEGRAND_RRM = 1 };

/* Component kinds (zodipy/number_density.py:408-420). */
enum {
  ZODI_CLOUD = 0,
  ZODI_BAND = 1,
  ZODI_RING = 2,
  ZODI_FEATURE = 3,
  ZODI_FAN = 4,
  ZODI_COMET = 5,
  ZODI_INTERSTELLAR = 6,
  ZODI_NARROW_BAND = 7,
  ZODI_BROAD_BAND = 8,
  ZODI_RING_RRM = 9,
  ZODI_FEATURE_RRM = 10
};

/* comp_flags bits */
#define ZODI_CF_CENTERED 1u   /* x_0 = y_0 = z_0 = 0: the component frame is heliocentric */
#define ZODI_CF_SCATTER 2u    /* albedo != 0: add scattered sunlight (zodipy/brightness.py:50) */
#define ZODI_CF_P_IS_4 4u     /* band exponent p == 4 exactly */

/* Per-component parameter slots, common part. */
enum { ZODI_P_X0 = 0, ZODI_P_Y0 = 1, ZODI_P_Z0 = 2,
       ZODI_P_NX = 3,   /* sin(Omega) sin(i) */
       ZODI_P_NY = 4,   /* cos(Omega) sin(i)  (enters with a minus sign) */
       ZODI_P_NZ = 5,   /* cos(i) */
       ZODI_P_K0 = 6,   /* first kind-specific parameter slot (6..12) */
       ZODI_P_D0 = 13,  /* first kind-specific derived slot (13..19) */
       ZODI_P_EMISSIVITY = 20, ZODI_P_ALBEDO = 21, ZODI_P_ONE_MINUS_ALBEDO = 22,
       ZODI_P_SCALE = 23, /* (1 - albedo) * emissivity */
       ZODI_P_SIN_I = 24, ZODI_P_COS_I = 25, ZODI_P_SIN_OMEGA = 26, ZODI_P_COS_OMEGA = 27 };

/* Shared parameter slots. */
enum { ZODI_S_C1 = 0, ZODI_S_C2 = 1, ZODI_S_C3 = 2, ZODI_S_SOLAR = 3,
       ZODI_S_PHASE_NORM = 4, ZODI_S_CALIBRATION = 5,
       ZODI_S_DD_DC3 = 6 /* dD/dC3 of the phase normalisation N = 1/(2 pi D) */ };

/*
 * Device model block: uploaded once per model construction or parameter update.
 * Components are listed in model order and partitioned into "groups": maximal
 * runs of consecutive components that share the line-of-sight range (cutoff
 * radii) and the dust temperature law, so position, temperature and Planck
 * lookup are evaluated once per group and node.
 *
 * Kind-specific slots, from ZODI_P_K0:
 *   CLOUD        n_0, alpha, beta, gamma, mu
 *   BAND         3*n_0, delta_zeta_rad, p, v, delta_r
 *   RING         n_0, R, sigma_r, sigma_z
 *   FEATURE      n_0, R, sigma_r, sigma_z, theta_rad, sigma_theta_rad
 *   FAN          Q, P, gamma, Z_0, R_outer
 *   COMET        gamma, Z_0, P, amp, R_inner, R_outer
 *   INTERSTELLAR amp
 *   NARROW_BAND  beta_nb, G, gamma, A, R_inner, R_outer
 *   BROAD_BAND   beta_bb, sigma_bb, gamma, A, R_inner, R_outer
 *   RING_RRM     n_0, R, sigma_r, sigma_z, A
 *   FEATURE_RRM  n_0, R, sigma_r, sigma_z, theta_rad, sigma_theta_rad, A
 * Derived slots, from ZODI_P_D0 (the kernels multiply instead of dividing):
 *   CLOUD        1/(2 mu), mu/2
 *   BAND         1/delta_zeta_rad, 1/v, 1/delta_r, then the plane normal (slots
 *                ZODI_P_NX..NZ) times 1/delta_zeta_rad, and 1/delta_r^2
 *   RING(_RRM)   1/sigma_r^2, 1/sigma_z, 1/sigma_r
 *   FEATURE(_RRM) 1/sigma_r^2, 1/sigma_z, 1/sigma_theta_rad^2, 1/sigma_r,
 *                1/sigma_theta_rad, cos(theta_rad), sin(theta_rad)
 *   FAN, COMET   1/Z_0
 *   NARROW_BAND  1/R_outer
 *   BROAD_BAND   1/sigma_bb, 1/R_outer
 */
typedef struct zodi_model {
  int32_t abi_version;
  int32_t integrand;
  int32_t n_comps;
  int32_t n_groups;
  int32_t n_nodes;
  int32_t n_temps;
  int32_t flags;
  int32_t reserved;
  int32_t comp_kind[ZODI_MAX_COMPS];
  uint32_t comp_flags[ZODI_MAX_COMPS];
  int32_t group_first[ZODI_MAX_GROUPS];
  int32_t group_count[ZODI_MAX_GROUPS];
  /* Parameter-vector positions for zodi_adjoint (see there). */
  int32_t grad_comp[ZODI_MAX_COMPS];    /* first of the component's init fields */
  int32_t grad_spec_a[ZODI_MAX_COMPS];  /* Kelsall: emissivity; RRM: T_0 of the component */
  int32_t grad_spec_b[ZODI_MAX_COMPS];  /* Kelsall: albedo; RRM: delta of the component */
  int32_t grad_shared[8];               /* Kelsall: T_0, delta, solar irradiance, C1, C2, C3;
                                           RRM: calibration */
  int32_t n_params;                     /* length of the parameter vector */
  int32_t reserved2;
  double group_r_in[ZODI_MAX_GROUPS];   /* start cutoff radius [AU] */
  double group_r_out[ZODI_MAX_GROUPS];  /* stop cutoff radius [AU] */
  double group_T0[ZODI_MAX_GROUPS];     /* T(R) = T0 R^-delta */
  double group_delta[ZODI_MAX_GROUPS];
  double shared[ZODI_SHARED_SLOTS];
  double comp[ZODI_MAX_COMPS][ZODI_COMP_SLOTS];
  double gl_x[ZODI_MAX_NODES];          /* Gauss-Legendre nodes on [-1, 1], ascending */
  double gl_w[ZODI_MAX_NODES];
  double tab_T[ZODI_N_TEMPS];           /* Planck table temperatures [K] */
  double tab_B[ZODI_N_TEMPS];           /* B_nu(T) [MJy/sr], delta or bandpass-averaged */
  double tab_slope[ZODI_N_TEMPS];       /* (B[j+1]-B[j])/(T[j+1]-T[j]) */
  double tab_inv_step;                  /* (n_temps-1)/(T[last]-T[0]) for the index guess */
  double pad_[7];
} zodi_model;

/* Observer addressing modes. */
enum {
  ZODI_OBS_PER_SEGMENT = 0, /* obs/earth are [3][n_seg]; ray r belongs to segment s with
                               seg_offsets[s] <= r < seg_offsets[s+1] */
  ZODI_OBS_PER_RAY = 1      /* obs/earth are [3][n_rays] (time-ordered data) */
};

/* Output modes. */
enum {
  ZODI_OUT_COMPS = 0,  /* out is [n_comps][n_rays] */
  ZODI_OUT_SUM = 1     /* out is [n_rays], components summed in model order */
};

/* Launch configuration (0 in any field selects the built-in default). */
typedef struct zodi_launch {
  int32_t block;          /* threads per block: 64, 128, 256, 512 */
  int32_t lanes_per_ray;  /* 1, 2, 4, 8: Gauss-Legendre nodes split over this many lanes */
                          /* (zodi_forward adds the node terms in node order, so its */
                          /* results do not depend on it; 0: chosen from n_rays)     */
  int32_t blocks_per_cu;  /* > 0: grid-stride launch of this many blocks per CU; 0 = one block per tile */
  int32_t variant;        /* 0 generic, 1 specialised layout if available */
} zodi_launch;

/* Library identity. */
int32_t zodi_abi_version(void);
int64_t zodi_model_size(void);             /* sizeof(zodi_model) */
const char* zodi_last_error(void);         /* thread-local message of the last failure */

/*
 * Every compute entry point takes the model block twice: `model_host` (host
 * memory) is validated and drives launch decisions; `model_dev` (device memory)
 * must hold the same bytes and is what the kernels read.
 *
 * zodi_forward — integrated emission per ray [MJy/sr].
 *
 * u          [3][n_rays] ecliptic unit vectors (SoA rows x, y, z)
 * seg_offsets[n_seg+1] ray offsets per observer (PER_SEGMENT), ignored for PER_RAY
 * obs, earth [3][n_obs] heliocentric ecliptic positions [AU]; n_obs = n_seg or n_rays
 * beyond     [n_seg] (PER_SEGMENT) or [1] (PER_RAY): bit 2g set when the observer lies
 *            outside group g's start sphere, bit 2g+1 for its stop sphere; such a range
 *            end is float64 epsilon (zodipy/line_of_sight.py:72-73: the test covers every
 *            observer of one reference call)
 * out        see ZODI_OUT_*
 * launch     NULL or a zodi_launch; every launch shape gives the same bits for a
 *            given variant (each ray's node terms are added in node order)
 */
zodi_status zodi_forward(const zodi_model* model_host, const zodi_model* model_dev,
                         const double* u, int64_t n_rays,
                         const int64_t* seg_offsets, int32_t n_seg, const double* obs,
                         const double* earth, int32_t obs_mode, const uint32_t* beyond,
                         double* out, int32_t out_mode, const zodi_launch* launch,
                         void* stream);

/*
 * zodi_line_of_sight — per-group start and stop distances [AU]: start/stop are
 * [n_groups][n_rays]. Same addressing as zodi_forward.
 */
zodi_status zodi_line_of_sight(const zodi_model* model_host, const zodi_model* model_dev,
                               const double* u, int64_t n_rays,
                               const int64_t* seg_offsets, int32_t n_seg, const double* obs,
                               int32_t obs_mode, const uint32_t* beyond, double* start,
                               double* stop, void* stream);

/*
 * zodi_number_density — component densities at heliocentric points:
 * X [3][n_points] -> out [n_comps][n_points]; earth [3] is the Earth position
 * used by the feature components.
 */
zodi_status zodi_number_density(const zodi_model* model_host, const zodi_model* model_dev,
                                const double* X, int64_t n_points,
                                const double* earth, double* out, void* stream);

/*
 * Pointing ingestion: ecliptic unit vectors out[3][n] (device memory) for
 * zodi_forward, replacing skycoord.transform_to(BarycentricMeanEcliptic) and
 * .cartesian (zodipy/model.py:247-251). rot is a 3x3 row-major rotation to the
 * mean ecliptic and equinox of J2000 in host memory, or NULL for none.
 *
 * zodi_pix2vec — centres of HEALPix RING pixels of resolution nside (what the
 * reference's HEALPix map examples build with healpy's pix2ang,
 * docs/examples/healpy_map.py:16): pix[n] (device memory), or, with pix NULL,
 * the pixels first .. first + n - 1. An index outside [0, 12 nside^2) gives NaN.
 *
 * zodi_lonlat_to_vec — (cos b cos l, cos b sin l, sin b) for longitudes lon[n]
 * and latitudes lat[n] (device memory), in degrees when degrees != 0, else radians.
 */
zodi_status zodi_pix2vec(int64_t nside, const int64_t* pix, int64_t first, int64_t n,
                         const double* rot, double* out, void* stream);

zodi_status zodi_lonlat_to_vec(const double* lon, const double* lat, int64_t n, int32_t degrees,
                               const double* rot, double* out, void* stream);

/*
 * zodi_spline_positions — heliocentric positions out[3][n] at times t[n] along a
 * cubic spline: the per-ray observer and Earth positions of time-ordered data,
 * which the reference interpolates on an hourly grid with scipy's CubicSpline
 * (zodipy/bodies.py:22-35). The spline is n_intervals records of 16 doubles in
 * device memory (16-byte aligned), one per interval k of ascending breakpoints:
 *   [0] x_k, [1..3] constant, [4..6] linear, [7..9] quadratic, [10..12] cubic
 *   coefficients of (x, y, z) in powers of (t - x_k), [13] x_{k+1}, [14..15] 0
 * (scipy's PPoly c[3-j][k][d] in slot 1 + 3 j + d). Intervals follow scipy
 * (extrapolating at both ends) and the sums round as scipy's do.
 */
zodi_status zodi_spline_positions(const double* records, int64_t n_intervals, const double* t,
                                  int64_t n, double* out, void* stream);

/*
 * zodi_max_radius — the largest distance to the Sun among per-ray observer
 * positions xyz[3][n] (device memory), written to out[0] (-1 when n = 0 or every
 * distance is NaN): the reference tests (r_obs > cutoff).any() over all
 * observers of one call (zodipy/line_of_sight.py:71-73), so per-ray observers
 * need this one number for their cutoff bits. out holds 1 + 1024 doubles of
 * device memory (out[1..] receives block partials).
 */
zodi_status zodi_max_radius(const double* xyz, int64_t n, double* out, void* stream);

/*
 * zodi_host_device_pointer — the device address of page-locked, mapped host
 * memory (hipHostMalloc, torch's pin_memory), so the kernels can read pointings
 * from, and write results to, host memory directly (zero-copy batches: one pass
 * over PCIe inside the kernel instead of separate copies). Fails with
 * ZODI_ERR_ARG for memory that is not page-locked host memory. No reference
 * counterpart: a helper for callers whose data stays in host memory.
 */
zodi_status zodi_host_device_pointer(const void* host, void** device);

/* Adjoint modes. */
enum {
  ZODI_ADJ_VJP_SUM = 0,    /* weights [n_rays] against the component sum */
  ZODI_ADJ_VJP_COMPS = 1,  /* weights [n_comps][n_rays] against each component */
  ZODI_ADJ_CHI2 = 2        /* evaluates I, chi2 = sum (I - target)^2 inv_var, and its
                              gradient (weights 2 (I - target) inv_var) */
};

/*
 * zodi_adjoint — gradient of sum_r sum_c weights_c,r I_c,r (or of chi2) with respect
 * to the parameter vector whose layout the model block's grad_* fields give.
 * No reference counterpart: the reference has no derivatives; its parameter
 * dictionary (zodipy/zodiacal_light_model.py:38-55, zodipy/model.py:303-333) fixes
 * the order of the component entries.
 *
 * Rays are addressed as in zodi_forward. grad is [n_params + 1] in device memory;
 * entry n_params is chi2 in ZODI_ADJ_CHI2 mode and 0 otherwise. inv_var is
 * [n_rays] or NULL (then inv_var_scalar applies to every ray). In ZODI_ADJ_CHI2
 * mode I [n_rays] is written to model_out, or to the workspace when model_out is
 * NULL. workspace holds zodi_adjoint_workspace() bytes of device memory for the
 * same n_rays, mode and launch. In ZODI_ADJ_CHI2 mode that size includes the
 * per-ray Jacobian when it fits in 72 GiB and launch allows one lane per ray:
 * one pass then writes the Jacobian and I (zodi_jacobian) and a second forms
 * chi2 and J^T g from them; with less workspace, or a larger problem, the
 * forward kernel writes I and the adjoint kernel recomputes the integrand with
 * the weights g. The two paths agree to rounding. launch may be NULL; otherwise
 * block must be 0 or 256, lanes_per_ray 0 (chosen from n_rays), 1, 2 or 4 (lanes
 * sharing one ray's nodes), blocks_per_cu 0..64 (0 = 2), and variant the
 * components evaluated per pass over a group's nodes: 0 (default: 4 in the
 * Jacobian pass, 2 otherwise), 1, 2 or 4 (four in the Jacobian pass, the sums of
 * up to two of them in LDS where registers do not hold all four; two in the
 * weighted pass).
 * Every component layout the block can describe is supported: the registry
 * layouts (the Kelsall cloud and three bands, optionally followed by the ring
 * and the feature; the RRM model in its registry grouping or one component per
 * group) run kernels specialised for them, any other combination of the eleven
 * kinds a generic kernel that takes one component per pass (variant is then
 * ignored). Mask edges of the RRM densities (R_inner, and R_outer where it only
 * masks) have derivative 0. The reduction order is fixed: equal inputs give
 * equal bits.
 */
int64_t zodi_adjoint_workspace(const zodi_model* model_host, int64_t n_rays, int32_t mode,
                               const zodi_launch* launch);

zodi_status zodi_adjoint(const zodi_model* model_host, const zodi_model* model_dev,
                         const double* u, int64_t n_rays, const int64_t* seg_offsets,
                         int32_t n_seg, const double* obs, const double* earth,
                         int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                         const double* weights, const double* target, const double* inv_var,
                         double inv_var_scalar, double* model_out, double* grad,
                         double* workspace, int64_t workspace_bytes, const zodi_launch* launch,
                         void* stream);

/*
 * zodi_table_vjp — gradient with respect to the block's Planck table B_k =
 * B(T_k), the bandpass-weighted blackbody emission that zodipy/blackbody.py:33-49
 * tabulates and zodipy/brightness.py:48 interpolates (np.interp at each node's
 * temperature). No reference counterpart (the reference has no derivatives):
 * with the effective-parameter gradient of zodi_adjoint it gives the gradient
 * with respect to the bandpass weights, in which both are linear. Arguments and
 * modes as in zodi_adjoint's recompute path; grad_table [ZODI_N_TEMPS + 1]
 * receives sum_r sum_c g_{c,r} dI_{c,r}/dB_k for k < n_temps (0 beyond), then
 * chi^2 (ZODI_ADJ_CHI2; else 0). A node between entries j and j + 1 at fraction f
 * contributes (1 - f) and f; a node at or past the last entry, or clamped below
 * the first, all of it to that entry. workspace: zodi_table_vjp_workspace bytes
 * (block partials, then I unless model_out is given). launch: block,
 * lanes_per_ray and blocks_per_cu as in zodi_adjoint (blocks_per_cu 0: as many
 * as the kernel keeps resident); variant is ignored. Any component layout. The
 * reduction order is fixed by the launch shape; within a wave, lanes that meet
 * in one entry add in hardware order.
 */
int64_t zodi_table_vjp_workspace(const zodi_model* model_host, int64_t n_rays,
                                 const zodi_launch* launch);

zodi_status zodi_table_vjp(const zodi_model* model_host, const zodi_model* model_dev,
                           const double* u, int64_t n_rays, const int64_t* seg_offsets,
                           int32_t n_seg, const double* obs, const double* earth,
                           int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                           const double* weights, const double* target, const double* inv_var,
                           double inv_var_scalar, double* model_out, double* grad_table,
                           double* workspace, int64_t workspace_bytes, const zodi_launch* launch,
                           void* stream);

/*
 * zodi_spectral_vjp — zodi_table_vjp together with the gradient with respect to
 * the effective spectral parameters of zodipy/brightness.py:21-83 (emissivity
 * and albedo of each Kelsall component, the solar irradiance and C1..C3; the
 * RRM calibration), in which the integrand is linear: everything a fit of the
 * bandpass weights or of the spectral tables (zodipy/unpack_model.py:140-173)
 * needs, in one pass over the nodes at about the cost of a forward pass, in
 * place of zodi_adjoint's full gradient. No reference counterpart (the reference
 * has no derivatives). grad [n_params + ZODI_N_TEMPS + 1]: entries 0..n_params-1
 * in zodi_adjoint's order, the effective spectral parameters' filled (they agree
 * with zodi_adjoint's to rounding) and every other entry 0; then the table
 * gradient as zodi_table_vjp writes it; then chi^2 (ZODI_ADJ_CHI2; else 0).
 * Other arguments, workspace (zodi_spectral_vjp_workspace bytes) and launch as
 * in zodi_table_vjp; the block's gradient layout must be valid (zodi_adjoint).
 */
int64_t zodi_spectral_vjp_workspace(const zodi_model* model_host, int64_t n_rays,
                                    const zodi_launch* launch);

zodi_status zodi_spectral_vjp(const zodi_model* model_host, const zodi_model* model_dev,
                              const double* u, int64_t n_rays, const int64_t* seg_offsets,
                              int32_t n_seg, const double* obs, const double* earth,
                              int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                              const double* weights, const double* target, const double* inv_var,
                              double inv_var_scalar, double* model_out, double* grad,
                              double* workspace, int64_t workspace_bytes,
                              const zodi_launch* launch, void* stream);

/*
 * zodi_adjoint_noalloc — zodi_adjoint without a workspace: no reference
 * counterpart (the reference has no derivatives). For callers that cannot
 * allocate device memory (code under HIP graph capture, tight fitting loops,
 * harnesses that forbid hipMalloc). It makes no allocation of any kind: the
 * adjoint kernel's block partials live in a 4 MiB array that the library holds
 * on each GPU it runs on until the process ends (at most 524288 doubles: launch
 * blocks x (n_params + 1), e.g. 512 blocks of 257 entries with the default
 * launch on a 256-CU GPU; larger launches return ZODI_ERR_TOO_LARGE). Every
 * entry the kernels read is written by the same call, so nothing carries over
 * between calls, but calls that overlap on one GPU (other streams, other host
 * threads) share the array and race. It always takes the recompute path: in
 * ZODI_ADJ_CHI2 mode model_out [n_rays] is required and receives I from the
 * forward kernel. Arguments, modes, layouts and launch as in zodi_adjoint; the
 * results equal zodi_adjoint's on its recompute path bit for bit.
 */
zodi_status zodi_adjoint_noalloc(const zodi_model* model_host, const zodi_model* model_dev,
                                 const double* u, int64_t n_rays, const int64_t* seg_offsets,
                                 int32_t n_seg, const double* obs, const double* earth,
                                 int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                                 const double* weights, const double* target,
                                 const double* inv_var, double inv_var_scalar, double* model_out,
                                 double* grad, const zodi_launch* launch, void* stream);

/* Targets of zodi_spectral_setup; component targets carry the component index
 * in the low byte (e.g. ZODI_SPEC_EMISSIVITY | 2). */
enum {
  ZODI_SPEC_EMISSIVITY = 0x100,
  ZODI_SPEC_ALBEDO = 0x200,
  ZODI_SPEC_SOLAR = 0x300,
  ZODI_SPEC_C1 = 0x301,
  ZODI_SPEC_C2 = 0x302,
  ZODI_SPEC_C3 = 0x303,
  ZODI_SPEC_CALIBRATION = 0x304 /* RRM; the MJy/AU -> Jy/cm factor goes into A */
};

/*
 * zodi_spectral_setup — writes effective spectral parameters and the Planck
 * table into a device model block, replacing zodipy/unpack_model.py:23-173
 * (interpolation of the spectral tables at the model's frequencies and the
 * bandpass trapezoid) and zodipy/blackbody.py:33-49 (the blackbody table) for
 * fits that move the tables. For each of n_tab targets, effective value =
 * sum_j A[t][j] tables[t][j] (A, tables: [n_tab][n_grid] device memory; A holds
 * the interpolation and trapezoid weights, the table's derivative), stored in
 * the target's slot; the derived slots (1 - albedo, (1 - albedo) emissivity,
 * the phase normalisation and dD/dC3) are then recomputed for the whole block.
 * With n_freq > 0 the table becomes B(T_k) = sum_i q[i] B_nu[i](T_k) in MJy/sr
 * at the block's 100 temperatures (nu_hz, q: [n_freq] device memory), with its
 * slopes. target is a host array. An albedo target needs the component packed
 * with its scattering branch (ZODI_CF_SCATTER). model_host supplies the block's
 * structure; its spectral values are not read. No reference counterpart for the
 * in-place update itself.
 */
zodi_status zodi_spectral_setup(const zodi_model* model_host, zodi_model* model_dev,
                                int32_t n_tab, int32_t n_grid, const double* A,
                                const double* tables, const int32_t* target, int32_t n_freq,
                                const double* nu_hz, const double* q, void* stream);

/*
 * zodi_jacobian — the per-ray Jacobian jac[p * ld_jac + r] = dI_r/dtheta_p
 * ([n_params][ld_jac], ld_jac >= n_rays) and the summed emission I_r
 * (model_out [n_rays]) in one pass, all in device memory. No reference
 * counterpart (the reference has no derivatives); the parameter order is that
 * of zodi_adjoint. For Gauss-Newton or Fisher-matrix work (J^T W J) and for the
 * chi2 mode of zodi_adjoint. Rays, layouts and launch as in zodi_adjoint, with
 * lanes_per_ray 0 or 1. I_r agrees with zodi_forward's to rounding.
 */
zodi_status zodi_jacobian(const zodi_model* model_host, const zodi_model* model_dev,
                          const double* u, int64_t n_rays, const int64_t* seg_offsets,
                          int32_t n_seg, const double* obs, const double* earth,
                          int32_t obs_mode, const uint32_t* beyond, double* model_out,
                          double* jac, int64_t ld_jac, const zodi_launch* launch, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ZODI_HIP_H */
