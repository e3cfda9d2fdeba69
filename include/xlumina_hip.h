/* C ABI of the MI355X light-propagation library (libxlumina_hip.so).
 *
 * Plain C: device pointers, sizes, float64 scalars, a hipStream_t passed as
 * void*.  Every call returns 0 on success or a positive status; the message of
 * the last failure on the calling thread is xl_last_error().  No call
 * allocates device memory: callers pass a workspace of at least
 * xl_*_workspace_bytes(...) bytes.  Calls are asynchronous on `stream`.
 *
 * Arrays: complex64 fields are interleaved (re, im) float pairs, row-major
 * [y][x], components `*_comp_stride` elements apart; masks are float32 [y][x];
 * grids x, y are float64 device arrays (the reference keeps x, y as float64).
 *
 * Each entry point names the reference function it replaces (paths relative to
 * the XLuminA repository root).
 */
#ifndef XLUMINA_HIP_H
#define XLUMINA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define XL_OK 0
#define XL_EINVAL 1     /* bad shape / argument (message says which)        */
#define XL_EHIP 2       /* HIP runtime error                                */
#define XL_EWORKSPACE 3 /* workspace smaller than *_workspace_bytes         */

/* Jones masks fused before a propagation (applied in order, up to two) */
#define XL_MASK_NONE 0
#define XL_MASK_SLM 1        /* u * exp(i m0)                    optical_elements.py:87-103  */
#define XL_MASK_SSLM 2       /* diag(exp(i m0), exp(i m1))       optical_elements.py:186-222 */
#define XL_MASK_LCD 3        /* retarder(eta=m0, theta=m1)       optical_elements.py:266-305 */
#define XL_MASK_LCD_SCALAR 4 /* retarder(eta=s0, theta=s1)       optical_elements.py:266-305 */
#define XL_MASK_LP 5         /* polariser alpha=m0, Ez -> 0      optical_elements.py:307-332 */
#define XL_MASK_AMP 6        /* real amplitudes: Ex*m0 (and Ey*m1 when m1 is set); with a
                                preceding XL_MASK_SSLM this is sSLM_with_amplitude
                                optical_elements.py:155-168, 224-264 */

/* construction of the propagated Ez from (Ex, Ey) (incoming Ez discarded) */
#define XL_EZ_NONE 0
#define XL_EZ_VRS 1  /* (Ex X + Ey Y)/r           vectorized_optics.py:258-261 */
#define XL_EZ_VCZT 2 /* (Ex X/r + Ey Y/r) z/r     vectorized_optics.py:341-344 */
#define XL_EZ_HNA 3  /* (Ex X + Ey Y)/rho         optical_elements.py:538      */

/* CZT variants */
#define XL_CZT_RS 0    /* F0 * K_y (E F) K_x^T * z dx dy lambda   wave_optics.py:333-357   */
#define XL_CZT_HNA 1   /* cte * K_y (RW-lens(E)) K_x^T            optical_elements.py:600-638 */
#define XL_CZT_PLAIN 2 /* K_y E K_x^T                             wave_optics.py:359-371   */

/* Storage forms of the mask parameters m0 / m1 (xl_mask_stage.forms: the
 * form of m0 in bits 0-2, of m1 in bits 3-5), so that a caller whose
 * parameters are float64 raw values -- the reference's optimisers keep them
 * so and map them with p * 2 pi - pi (optical_elements.py:1556-1559) -- passes
 * them without a cast, a rescaling or an expansion to a per-pixel array:
 *   0                                 float32 per pixel (const float*)
 *   XL_PF_F64                         float64 per pixel (const double*)
 *   XL_PF_F64 | XL_PF_SCALAR          one float64 on the device for every pixel
 *   ... | XL_PF_UNIT_ANGLE            the value used is 2 pi p - pi of the stored p
 * The float64 forms need batch <= 1.  xl_pre_backward writes the gradient of
 * the stored p: float32 per pixel for form 0, float64 per pixel for XL_PF_F64
 * (gparam then points to doubles), and for a scalar form the sum over the
 * pixels at element 0 of a buffer of xl_pre_scalar_grad_elems(rows, cols)
 * doubles. */
#define XL_PF_F64 1
#define XL_PF_SCALAR 2
#define XL_PF_UNIT_ANGLE 4
#define XL_RS_SPEC_PREBUILT 4   /* spec_grads bit: the derivative spectra are already in `spec` */
#define XL_RS_SPEC_PARTIALS 8   /* spec_grads bit of spec_mode 2: the per-unit partial sums stay in `spec`
                                   (xl_rs_spec_partials) and grad_out is not written */

typedef struct xl_mask_stage {
  int32_t kind;
  int32_t forms;   /* storage forms of m0 (bits 0-2) and m1 (bits 3-5), XL_PF_*  */
  const float* m0; /* parameter 0 (phase / eta / alpha) or NULL              */
  const float* m1; /* parameter 1 (phi / theta) or NULL                      */
  double s0, s1;   /* constant eta, theta for XL_MASK_LCD_SCALAR           */
} xl_mask_stage;

typedef struct xl_rs_params {
  const void* in;          /* nin complex64 components, n x n           */
  int64_t in_comp_stride;
  int32_t nin;             /* 1 (scalar) or 2/3 (Ex, Ey[, Ez])          */
  int32_t ez_mode;         /* XL_EZ_NONE or XL_EZ_VRS                   */
  void* out;               /* ncomp complex64 components, n x n         */
  int64_t out_comp_stride;
  float* intensity;        /* optional n x n sum_c |out_c|^2, or NULL   */
  int32_t n;               /* grid side (square, as the reference)      */
  int32_t adjoint;         /* 1: apply the adjoint operator (conj kernel) */
  const double* x;         /* n float64 samples                         */
  const double* y;
  double z;                /* propagation distance (microns)            */
  double k;                /* 2 pi / lambda                              */
  double dx, dy;           /* x[1]-x[0], y[1]-y[0] (wave_optics.py:269-270) */
  xl_mask_stage mask[2];
  const double* z_dev;     /* optional device copy of z read by the kernels instead of `z`
                              (lets a captured HIP graph replay with a new distance) */
  int32_t kernel_dz;       /* 1: convolve with dH/dz instead of H (derivative of the
                              propagation with respect to z, for gradients in z); 2: with
                              dH/dk, k = 2 pi / lambda (for gradients in the wavelength) */
  int32_t hhat_mode;       /* 0: the kernel spectrum is built in the workspace; 1: built into `hhat`
                              and kept there; 2: read from `hhat` as a mode-1 call with the same n,
                              x, y, z, k, dx, dy left it, skipping the two kernel passes (an adjoint
                              call conjugates it on read) -- a gradient step's adjoint propagation
                              reuses its forward's spectrum.  Modes 1 and 2 need kernel_dz = 0. */
  void* hhat;              /* modes 1, 2: xl_rs_hhat_bytes(n) bytes of device memory */
  /* Batch of independent problems of this geometry run by the same launches
   * (batch <= 1: one problem).  Member b reads in + b*in_batch_stride and the
   * mask parameters + b*mask_batch_stride (0: shared by the batch), writes
   * out + b*out_batch_stride and intensity + b*intensity_batch_stride, and,
   * when z_dev is set, propagates over z_dev[b].  The workspace holds one
   * block per member (xl_rs_batch_workspace_bytes); with hhat_mode 1 / 2,
   * hhat holds one spectrum per member (batch * xl_rs_hhat_bytes(n) bytes). */
  int32_t batch;
  int32_t reserved2;
  int64_t in_batch_stride;        /* complex elements (0: one input shared by the batch) */
  int64_t out_batch_stride;       /* complex elements, >= ncomp n^2 */
  int64_t intensity_batch_stride; /* float elements, >= n^2 when intensity is set */
  int64_t mask_batch_stride;      /* float elements (0: shared mask parameters) */
  /* Gradients in z and in k = 2 pi / lambda without a derivative propagation
   * (one problem, kernel_dz = 0): by Parseval, Re<g, conv(E', dH)> =
   * sum_c sum_k Re(conj(G_c(k)) dH(k) U_c(k)) over the P x P spectra, where U_c
   * is the spectrum of the padded, masked input component c and G_c that of
   * the output cotangent.
   *   spec_mode 1 (forward call): U_c is kept in `spec`
   *     (xl_rs_spec_bytes(n, ncomp) bytes of device memory); the derivative
   *     spectra named by spec_grads (bit 0: dH/dz, bit 1: dH/dk; 0: none) are
   *     built there too, from the kernel rows the call evaluates for H;
   *   spec_mode 2 (the adjoint call of the same geometry and distance, its
   *     input the output cotangent): the derivative spectra named by
   *     spec_grads are built in `spec` -- or, with XL_RS_SPEC_PREBUILT set,
   *     taken from there as the mode-1 call with those bits left them -- and
   *     grad_out[0] = Re<g, conv(E', dH/dz)>, grad_out[1] = Re<g, conv(E', dH/dk)>
   *     (float64, fixed summation order) are written for the bits set.
   * The Ez_in construction's own dependence on z is not included (xl_dz_dot
   * with ncomp = 0 adds it). */
  int32_t spec_mode;       /* 0, 1 or 2 */
  int32_t spec_grads;      /* bit 0 z, bit 1 k; spec_mode 2 may add XL_RS_SPEC_PREBUILT */
  void* spec;
  double* grad_out;        /* spec_mode 2: 2 float64 on the device */
  /* Optional: the quality factor of the distance propagated (z_dev[b] or z),
   * (sqrt(wl^2 + rmax^2 + 2 wl sqrt(rmax^2 + z^2)) - rmax) / dr_real
   * (wave_optics.py:187-191), written by the launch itself into
   * quality_factor[b] (one float64 per member), so that a propagation to a
   * distance held on the device reports it without a launch of its own. */
  double* quality_factor;
  double qf_rmax, qf_dr_real, qf_wavelength;
  /* Optional: the derivative spectra of spec_mode 1 / 2 live in this buffer
   * instead of `spec`: the spec buffer of another call of the same n, ncomp,
   * x, y, z and k whose spec_grads covered these bits and which built them.  A
   * spec_mode 1 call then builds none (pair it with hhat_mode 2 on that
   * call's hhat to share its kernel spectrum too); a spec_mode 2 call needs
   * XL_RS_SPEC_PREBUILT with it.  Two propagations over one distance share
   * their kernel and derivative spectra this way. */
  void* spec_dg;
} xl_rs_params;

typedef struct xl_czt_params {
  const void* in;          /* nin components, rows x cols of the input plane */
  int64_t in_comp_stride;
  int32_t nin;
  int32_t ncomp;           /* components propagated (3 when Ez is built)     */
  void* out;
  int64_t out_comp_stride;
  int32_t ny_in, nx_in;    /* input plane (forward)                           */
  int32_t my, mx;          /* output plane (forward)                          */
  int32_t mode;            /* XL_CZT_*                                        */
  int32_t adjoint;         /* 1: in = (my, mx) cotangent, out = (ny_in, nx_in) */
  int32_t ez_mode;         /* XL_EZ_NONE / XL_EZ_VCZT (XL_CZT_HNA builds its own) */
  int32_t reserved;
  const double* x;         /* input-plane grid (nx_in, ny_in)                  */
  const double* y;
  const double* xout;      /* output-plane grid (mx, my)                       */
  const double* yout;
  double z, wavelength, k;
  double dm;               /* Dm of the Bluestein frequency window             */
  double fy1, fy2, fx1, fx2;
  double cte_re, cte_im;   /* constant output factor (forward)                 */
  double hna_radius, hna_f;
  xl_mask_stage mask[2];
  /* Batch of independent problems sharing this geometry, run by the same
   * launches (Bluestein form only).  batch <= 1: one problem. */
  int32_t batch;
  int32_t tables_mode;         /* Bluestein form: 0 stages the axis tables (chirps and their spectra) in
                                  this call; 1 stages them into the workspace and returns (no other
                                  launch, in / out untouched); 2 uses the tables a mode-1 call with the
                                  same parameters staged in this workspace -- so a caller can stage
                                  them on another stream while the input is still being produced */
  int64_t in_batch_stride;    /* complex elements between problems' inputs (0: one shared input) */
  int64_t out_batch_stride;   /* complex elements between problems' outputs */
  int64_t mask_batch_stride;  /* float elements between problems' per-pixel mask parameters (0: shared) */
} xl_czt_params;

typedef struct xl_pre_params {
  /* forward pre-op: masks on (Ex, Ey) then XL_EZ_* / high-NA lens map;
   * backward: cotangent of the pre-op output -> cotangents of Ex, Ey and of
   * the per-pixel mask parameters. */
  const void* in;          /* Ex, Ey (complex64, n_rows x n_cols)              */
  int64_t in_comp_stride;
  int32_t nin;
  int32_t ncomp;           /* components of the pre-op output                 */
  const void* gout;        /* backward: cotangent of the pre-op output         */
  void* out;               /* forward: pre-op output; backward: grad of inputs */
  int64_t comp_stride;
  float* gparam[2][2];     /* backward: grads of mask[s].m0 / m1 in their storage forms, or NULL */
  int32_t rows, cols;
  int32_t ez_mode;         /* XL_EZ_*                                          */
  int32_t hna;             /* 1: high-NA lens map (optical_elements.py:555-594) */
  const double* x;
  const double* y;
  double z;
  double hna_radius, hna_f;
  xl_mask_stage mask[2];
  const double* z_dev;     /* optional device copy of z (see xl_rs_params); with batch > 1 one
                              distance per member, z_dev[b] */
  int32_t batch;           /* independent problems per launch (<= 1: one) */
  int32_t out_planes;      /* backward: planes of `out`; planes nin .. out_planes-1 (the components
                              of the source the pre-op does not read) receive zeros (<= nin: none) */
  int64_t in_batch_stride;     /* complex elements (0: one input shared by the batch) */
  int64_t gout_batch_stride;   /* complex elements */
  int64_t out_batch_stride;    /* complex elements */
  int64_t mask_batch_stride;   /* float elements (0: shared mask parameters) */
  int64_t gparam_batch_stride; /* float elements */
  /* backward with XL_EZ_VRS (one problem), optional: xl_pre_scalar_grad_elems(rows, cols)
   * float64 whose element 0 receives sum Re(conj(g_Ez) Ex') X w + Re(conj(g_Ez) Ey') Y w,
   * w = -z / (X^2 + Y^2 + z^2)^{3/2}: the distance dependence of the constructed
   * Ez_in = (Ex' X + Ey' Y)/r (vectorized_optics.py:258-261) through its cotangent, the
   * term xl_dz_dot forms from `pre` -- here without a second pass over the field. */
  double* ez_dz;
  const double* ez_dz_base; /* optional: float64 on the device added to ez_dz[0] (the
                               propagation's own term, so element 0 holds the whole gradient) */
  int32_t ez_dz_base_n;     /* values at ez_dz_base, summed in a fixed order (0: one) -- e.g. the
                               per-unit partial sums an XL_RS_SPEC_PARTIALS adjoint left in its spec
                               buffer (xl_rs_spec_partials) */
  int32_t reserved3;
} xl_pre_params;

const char* xl_last_error(void);
int xl_abi_version(void);
int xl_max_fft_length(void);
/* Builds the float64-generated twiddle table on the current device (called
 * once per device before any other call; later calls are no-ops). */
int xl_init(void);

/* RS_propagation_jit (wave_optics.py:281-289), VRS_propagation_jit
 * (vectorized_optics.py:364-373) with Jones masks / Ez construction fused. */
size_t xl_rs_workspace_bytes(int32_t n, int32_t ncomp);
size_t xl_rs_batch_workspace_bytes(int32_t n, int32_t ncomp, int32_t batch);
size_t xl_rs_hhat_bytes(int32_t n);
/* Bytes of the `spec` buffer of spec_mode 1 / 2 (the input spectra, two
 * derivative spectra and per-column partial sums); 0 for bad arguments.
 * The gradients in z / lambda of RS_propagation (wave_optics.py:281-289
 * under jax.grad) without a second propagation. */
size_t xl_rs_spec_bytes(int32_t n, int32_t ncomp);
/* Byte offset, from the 256-byte-aligned start of a spec buffer, of the
 * float64 per-unit partial sums of derivative i (0: z, 1: k) that a spec_mode 2
 * call writes; *count receives their number.  -1 for bad arguments. */
int64_t xl_rs_spec_partials(int32_t n, int32_t ncomp, int32_t i, int32_t* count);
int xl_rs_propagate(const xl_rs_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference.  xl_rs_propagate without a workspace
 * argument, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * kernels use a library-owned scratch array of XL_RS_NOALLOC_SCRATCH_BYTES
 * bytes, one per device, held from the device's xl_init until the process
 * ends.  Grids above XL_RS_NOALLOC_MAX_N per side return XL_EINVAL.  The call
 * allocates nothing.  Every call on one device shares that device's array:
 * calls that overlap on one device (other streams, other host threads) race
 * and must be serialised by the caller.  The hhat buffer of hhat_mode 1 / 2
 * is still the caller's. */
#define XL_RS_NOALLOC_MAX_N 1024
#define XL_RS_NOALLOC_SCRATCH_BYTES 67125504 /* xl_rs_workspace_bytes(1024, 3) */
int xl_rs_propagate_noalloc(const xl_rs_params* p, void* stream);

/* CZT_jit (wave_optics.py:333-357), VCZT_jit (vectorized_optics.py:375-384),
 * CZT_for_high_NA_jit (wave_optics.py:359-371), VCZT_objective_lens
 * (optical_elements.py:600-638); Bluestein FP32 form.  Returns XL_EINVAL for
 * the shapes the reference cannot run (m + M - 1 a power of two, M = 1). */
size_t xl_czt_workspace_bytes(const xl_czt_params* p);
int xl_czt(const xl_czt_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* Same operator as xl_czt (same params, forward and adjoint) as a dense
 * zoom-DFT on MFMA: K_y G K_x^T with the kernels generated in float64 phase
 * and split into bf16 hi + lo planes, FP32 accumulation (12 bf16 MFMAs per
 * complex MAC); relative error ~1e-5.  Pays off for small output windows. */
size_t xl_czt_gemm_workspace_bytes(const xl_czt_params* p);
int xl_czt_gemm(const xl_czt_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* Standalone pre-ops and their adjoint: SLM / sSLM / LCD / linear_polarizer
 * (optical_elements.py:87-332), high_NA_objective_lens (:515-594). */
int xl_pre_forward(const xl_pre_params* p, void* stream);
int xl_pre_backward(const xl_pre_params* p, void* stream);
/* doubles of the gradient buffer of a scalar-form mask parameter in
 * xl_pre_backward over rows x cols pixels: the sum at element 0, one partial
 * sum per launch block after it (see XL_PF_SCALAR). */
size_t xl_pre_scalar_grad_elems(int32_t rows, int32_t cols);

/* MSE_Intensity (loss_functions.py:137-149) of a batch of detector fields
 * against one target intensity |T|^2, and its cotangent with respect to the
 * field: loss_b = scale * sum_p (|E_bp|^2 - T_p)^2 / npix,
 * grad_bp = scale * 4 (|E_bp|^2 - T_p) E_bp / npix (dL/dRe + i dL/dIm).
 * grad may alias field.  Deterministic two-stage reduction. */
typedef struct xl_mse_params {
  const void* field;           /* batch x npix complex64 */
  int64_t field_batch_stride;
  const float* target;         /* npix float32 target intensity */
  void* grad;                  /* batch x npix complex64, or NULL */
  int64_t grad_batch_stride;
  float* loss;                 /* batch float32 losses, or NULL */
  int64_t npix;
  int32_t batch;
  float scale;
} xl_mse_params;
size_t xl_mse_workspace_bytes(int64_t npix, int32_t batch);
int xl_mse_intensity(const xl_mse_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* One Adam / AdamW update (optax.adam; optax.adamw as used by
 * experiments/four_f_optimizer.py:112) of n float32 parameters in place:
 * m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, t = count + 1,
 * p -= lr ((m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps) + weight_decay p);
 * then count += 1.  count lives in device memory so that a captured HIP
 * graph of several steps can be replayed. */
typedef struct xl_adam_params {
  float* param;
  const float* grad;
  float* m;
  float* v;
  int64_t n;
  int32_t* count;
  float lr, b1, b2, eps, weight_decay;
  int32_t reserved;
} xl_adam_params;
int xl_adam_step(const xl_adam_params* p, void* stream);

/* Field losses against a complex64 target field (loss_functions.py:107-149),
 * for a batch of fields: loss_b = scale * sum_p d_bp^2 / npix with
 *   XL_LOSS_INTENSITY  d = |E|^2 - |T|^2          (MSE_Intensity :137-149)
 *   XL_LOSS_AMPLITUDE  d = |E| - |T|              (MSE_Amplitude :107-120)
 *   XL_LOSS_PHASE      d = angle(E) - angle(T)    (MSE_Phase     :122-135)
 * The batch covers vMSE_* (:65-105, members = the Ex, Ey, Ez components) and
 * mean_batch_MSE_Intensity (:51-63).  grad (optional) is the cotangent of the
 * field, dL/dRe + i dL/dIm: 4 s d E, 2 s d E/|E|, 2 s d i E/|E|^2 (s = scale /
 * npix; the last two are 0 where E = 0).  The target is a constant.
 * Deterministic two-stage reduction. */
#define XL_LOSS_INTENSITY 0
#define XL_LOSS_AMPLITUDE 1
#define XL_LOSS_PHASE 2
typedef struct xl_field_loss_params {
  const void* field;            /* batch x npix complex64 */
  int64_t field_batch_stride;
  const void* target;           /* complex64 target fields */
  int64_t target_batch_stride;  /* 0: one target for every member */
  void* grad;                   /* batch x npix complex64, or NULL; may alias field */
  int64_t grad_batch_stride;
  float* loss;                  /* batch float32 losses, or NULL */
  int64_t npix;
  int32_t batch;
  int32_t mode;                 /* XL_LOSS_* */
  float scale;
  int32_t reserved;
} xl_field_loss_params;
size_t xl_field_loss_workspace_bytes(int64_t npix, int32_t batch);
int xl_field_loss(const xl_field_loss_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* small_area_hybrid (loss_functions.py:24-40) of a batch of detected
 * intensities D_b (vectorized_loss_hybrid :42-49): with S = sum D,
 * I = D / (S + 1e-8) and mask m = [I > 0.7 max I],
 * loss_b = scale * sum m / (sum m I + 1e-8).  grad (optional) = dloss/dD with
 * the mask held constant (as jnp.where differentiates):
 * -scale * (sum m) / (sum m I + 1e-8)^2 * (m_p - sum m I) / (S + 1e-8). */
typedef struct xl_small_area_params {
  const float* intensity;      /* batch x npix float32 */
  int64_t batch_stride;
  float* grad;                 /* batch x npix float32, or NULL */
  int64_t grad_batch_stride;
  float* loss;                 /* batch float32 */
  int64_t npix;
  int32_t batch;
  float scale;
} xl_small_area_params;
size_t xl_small_area_workspace_bytes(int64_t npix, int32_t batch);
int xl_small_area(const xl_small_area_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* Two-port beam splitter on (Ex, Ey) with complex coefficients:
 *   c = ca a + cb b,  d = da a + db b   (b = NULL: single input)
 * which is BS_symmetric (optical_elements.py:334-392: ca = db = R, cb = da = iT
 * with the 1% loss folded into R, T), BS_symmetric_SI (:394-460: b = NULL,
 * port 1 swaps c and d) and BS (:462-513: ca = db = sqrt T,
 * cb = sqrt R e^{i phase}, da = -sqrt R e^{-i phase}).  The adjoint is the
 * same call with (a, b) = cotangents of (c, d) and coefficients
 * (conj ca, conj da, conj cb, conj db).  c or d may be NULL; with
 * zero_ez = 1 component 2 of each output is set to 0 (the reference builds c
 * and d as new lights, Ez = 0). */
#define XL_BS_THETA_RAW 4   /* added to theta_mode 1..3: theta is a raw angle parameter */

typedef struct xl_bs_params {
  const void* a;               /* 2 components (Ex, Ey) of npix complex64 */
  const void* b;               /* or NULL */
  int64_t a_comp_stride, b_comp_stride;
  void* c;                     /* or NULL */
  void* d;                     /* or NULL */
  int64_t c_comp_stride, d_comp_stride;
  float ca[2], cb[2], da[2], db[2];   /* (re, im) */
  int64_t npix;
  int32_t zero_ez;
  int32_t theta_mode;          /* 0: the coefficients above (or coef); 1: BS_symmetric (R, iT, iT, R),
                                  2: BS_symmetric_SI port 0 (R, 0, iT, 0), 3: port 1 (iT, 0, R, 0), with
                                  R = |sin t| - 0.01 |cos t|, T = 0.99 |cos t| formed on the device from
                                  theta (optical_elements.py:376-383, 394-460); + XL_BS_THETA_RAW: theta
                                  holds the raw parameter p of t = 2 pi p - pi (optical_elements.py:1590) */
  const float* coef;           /* optional device array (ca, cb, da, db as re, im pairs: 8 floats) read
                                  by the kernel instead of the values above (no host round trip when
                                  the coefficients are computed on the device) */
  const double* theta;         /* theta_mode > 0: the splitter angle t (one float64 on the device) */
  int32_t adjoint;             /* theta_mode > 0: apply the conjugate transpose, i.e. the cotangents of
                                  (a, b) from those of (c, d) passed as a and b */
  int32_t reserved;
} xl_bs_params;
int xl_beam_splitter(const xl_bs_params* p, void* stream);

/* Coefficient cotangents of a beam splitter: dots = the four complex sums
 * over both components and all pixels of conj(gc) a, conj(gc) b, conj(gd) a,
 * conj(gd) b as 8 float32 (re, im); a NULL gc, gd or b contributes 0.
 * For a real parameter t of the coefficients, dL/dt = Re sum_k kappa_k'(t) S_k; for the
 * splitters of theta_mode 1-3 the kernel writes that derivative itself. */
typedef struct xl_bs_dots_params {
  const void* gc;
  const void* gd;
  int64_t gc_comp_stride, gd_comp_stride;
  const void* a;
  const void* b;
  int64_t a_comp_stride, b_comp_stride;
  float* dots;                 /* 8 float32 in device memory, or NULL */
  int64_t npix;
  const double* theta;         /* with theta_mode > 0 (as in xl_bs_params): the splitter angle */
  int32_t theta_mode;
  int32_t reserved;
  double* dtheta;              /* with theta_mode > 0: dL/dt = Re sum_k S_k dkappa_k/dt (one float64);
                                  with XL_BS_THETA_RAW, dL/dp = 2 pi dL/dt */
  /* Optional, with theta_mode > 0: the input cotangents in the same pass (the
   * xl_beam_splitter adjoint), ga = conj(ca) gc + conj(da) gd and
   * gb = conj(cb) gc + conj(db) gd on components 0, 1 (an absent gc / gd
   * counts as 0); planes 2 .. g_planes-1 of ga / gb receive zeros. */
  void* ga;
  void* gb;
  int64_t ga_comp_stride, gb_comp_stride;
  int32_t g_planes;
  int32_t reserved2;
} xl_bs_dots_params;
size_t xl_bs_dots_workspace_bytes(int64_t npix);
int xl_bs_dots(const xl_bs_dots_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* Second moment of a focal spot, L = sum_p I_p r2_p / sum_p I_p with
 * I = sum_c |E_c|^2 over ncomp components (the high-NA focusing loss of the
 * optimisation benchmark; no counterpart function in the reference, whose
 * optimisers build losses from jnp reductions), in one fused pass with a
 * fixed-order reduction.  forward writes stats = (sum I, L) and loss (both
 * device floats; loss may be NULL); backward writes the field cotangent
 * grad_c = g 2 (r2 - L) E_c / sum I, with g read from device memory
 * (`gscale`, the incoming gradient of L; NULL means 1), so neither direction
 * synchronises with the host. */
typedef struct xl_moment_params {
  const void* field;           /* ncomp x npix complex64 */
  int64_t comp_stride;
  int32_t ncomp;
  int32_t reserved;
  const float* r2;             /* npix weights (u^2 + v^2 of the focal window) */
  int64_t npix;
  float* stats;                /* 2 floats: sum I, L */
  float* loss;                 /* forward: L (device float), or NULL */
  const float* gscale;         /* backward: incoming gradient of L (device float), or NULL */
  void* grad;                  /* backward: ncomp x npix complex64 (comp_stride apart) */
} xl_moment_params;
size_t xl_moment_workspace_bytes(int64_t npix);
int xl_moment_forward(const xl_moment_params* p, void* workspace, size_t workspace_bytes, void* stream);
int xl_moment_backward(const xl_moment_params* p, void* stream);

/* Quality factor of a propagation to a distance held on the device
 * (wave_optics.py:187-191): out = (sqrt(wl^2 + rmax^2 + 2 wl sqrt(rmax^2 + z^2))
 * - rmax) / dr_real with z = z_dev[0]; rmax and dr_real as the reference
 * derives them from the grid.  One float64 written on the stream. */
int xl_quality_factor(const double* z_dev, double rmax, double dr_real, double wavelength, double* out,
                      void* stream);

/* Distance gradient of one RS / VRS propagation (the reduction half of
 * rs_propagation_dz, wave_optics.py:173-196 / vectorized_optics.py:244-284
 * differentiated in z): out[0] = sum_c Re(conj(d_c) g_c) over `ncomp` planes
 * of the dz-propagated field d and the output cotangent g, plus, when `pre`
 * is set (VRS), sum Re(conj(gez) Ex') X w + Re(conj(gez) Ey') Y w with
 * w = -z / (X^2 + Y^2 + z^2)^{3/2}: the z-dependence of the constructed
 * Ez_in = (Ex' X + Ey' Y)/r (vectorized_optics.py:258-261), where pre holds
 * (Ex', Ey') and gez the cotangent of Ez_in.  float64 result on the device,
 * summed in a fixed order.  ncomp = 0 (d and g unused) with `pre` set gives
 * the Ez_in term alone. */
typedef struct xl_dz_dot_params {
  const void* d;               /* ncomp x rows x cols complex64, comp_stride apart */
  const void* g;               /* same layout as d */
  int64_t comp_stride;
  int32_t ncomp;
  int32_t rows;
  int32_t cols;
  int32_t reserved;
  const void* pre;             /* 2 planes complex64 (pre_comp_stride apart), or NULL */
  int64_t pre_comp_stride;
  const void* gez;             /* rows x cols complex64 (with pre) */
  const double* x;             /* cols grid (with pre) */
  const double* y;             /* rows grid (with pre) */
  double z;
  const double* z_dev;         /* device copy of z read instead of `z`, or NULL */
  double* out;                 /* one float64 on the device */
} xl_dz_dot_params;
size_t xl_dz_dot_workspace_bytes(void);
int xl_dz_dot(const xl_dz_dot_params* p, void* workspace, size_t workspace_bytes, void* stream);

/* The adjoint of one XL_MASK_SLM stage on one component (xl_pre_backward
 * with nin = ncomp = 1) fused with the Adam / AdamW step of those phases: per
 * pixel dL/dphi = Re(conj(g) i E e^{i phi}) stays in registers and the phases
 * (mask[0].m0, which must be adam->param), m and v (laid out like the phases:
 * same batch stride, mask_batch_stride) take xl_adam_step's update with
 * t = count + 1.  count is NOT advanced: call xl_adam_tick once after the
 * last update of a training step.  adam->grad is ignored; gparam[0][0], if
 * set, still receives the phase cotangent and out the field cotangent.
 * What optax.adam(w) after jax.grad of an SLM (optical_elements.py:87-103,
 * experiments/four_f_optimizer.py:112) computes, without a gradient array in
 * HBM. */
int xl_slm_backward_adam(const xl_pre_params* p, const xl_adam_params* adam, void* stream);
/* count += 1 on the stream (one thread). */
int xl_adam_tick(int32_t* count, void* stream);

/* transfer_function_RS (wave_optics.py:291-297) on meshgrid(x, y): out is
 * ny x nx complex64. */
int xl_transfer_function_rs(void* out, const double* x, const double* y, int32_t nx, int32_t ny, double z,
                            double k, void* stream);

/* Seeded uniform phase masks (no counterpart in the reference, whose phases
 * come from the caller; used for synthetic request streams): for each member
 * b < batch with enable[b] != 0 (enable NULL: every member),
 * out[b * batch_stride + i] = 2 pi u - pi for i < count, u in [0, 1) from a
 * counter-based hash (splitmix64) of seeds[b] and i.  Seeds and flags are
 * device arrays, so a captured HIP graph regenerates a batch from new seeds. */
int xl_uniform_phases(float* out, const uint64_t* seeds, const int32_t* enable, int64_t count, int32_t batch,
                      int64_t batch_stride, void* stream);

/* Opt-in per-kernel timing: HIP events recorded on the launch stream around
 * every kernel launch of the calls that follow (capacity = launches kept).
 * xl_timing_collect waits on them, writes per-kernel totals (up to
 * max_entries) and returns the number of distinct kernels (negative on error);
 * it empties the record list.  Keep it off during graph capture. */
typedef struct xl_kernel_time {
  char name[32];
  int64_t launches;
  double total_ms;
} xl_kernel_time;
int xl_timing_enable(int32_t enable, int32_t capacity);
int xl_timing_collect(xl_kernel_time* out, int32_t max_entries);

#ifdef __cplusplus
}
#endif
#endif /* XLUMINA_HIP_H */
