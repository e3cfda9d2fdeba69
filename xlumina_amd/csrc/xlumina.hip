// C ABI of libxlumina_hip.so (declared in include/xlumina_hip.h).
//
// One translation unit: the kernels live in the headers included here, the
// entry points below validate shapes on the host (float64 scalar set-up, as the
// reference's build_grid / build_CZT_grid do in NumPy), carve the caller's
// workspace and launch on the caller's stream.  Nothing here allocates device
// memory or synchronises the stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/xlumina_hip.h"
#include "czt.h"
#include "opt.h"
#include "gemm.h"
#include "fft.h"
#include "masks.h"
#include "mix.h"
#include "pre.h"
#include "rng.h"
#include "rs.h"
#include "timing.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(XL_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return XL_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_init_mu;
bool g_init_done[kMaxDevices] = {};

// Scratch of xl_rs_propagate_noalloc: one copy per device in the code
// object's zero-initialised data.  In host code `scratch` names a host-side
// stand-in, so the device address of each device's copy is looked up once at
// initialisation (under g_init_mu) and kept in g_scratch.
static __device__ __align__(256) char scratch[XL_RS_NOALLOC_SCRATCH_BYTES];
void* g_scratch[kMaxDevices] = {};

int ensure_init() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(XL_EHIP, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_init_done[dev]) return XL_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(nullptr, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
    return fail(XL_EINVAL, "xl_init must run before graph capture");
  hipLaunchKernelGGL(xl::xl_fft_init_kernel, dim3((xl::TWN + 255) / 256), dim3(256), 0, nullptr);
  int rc = hip_check(hipGetLastError(), "twiddle init launch");
  if (rc) return rc;
  rc = hip_check(hipDeviceSynchronize(), "twiddle init");
  if (rc) return rc;
  // launch attributes are per device: set them here, once, under the lock
  rc = hip_check(xl::rs_set_attributes(), "RS launch attributes");
  if (rc) return rc;
  rc = hip_check(xl::czt_set_attributes(), "CZT launch attributes");
  if (rc) return rc;
  rc = hip_check(hipGetSymbolAddress(&g_scratch[dev], HIP_SYMBOL(scratch)), "scratch address");
  if (rc) return rc;
  g_init_done[dev] = true;
  return XL_OK;
}

bool device_ready() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return false;
  std::lock_guard<std::mutex> lk(g_init_mu);
  return g_init_done[dev];
}

int next_pow2(int64_t v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

xl::MaskArgs to_mask(const xl_mask_stage* m) {
  xl::MaskArgs a;
  for (int i = 0; i < 2; ++i) {
    a.st[i].kind = m[i].kind;
    a.st[i].pf = m[i].forms;
    a.st[i].m0 = m[i].m0;
    a.st[i].m1 = m[i].m1;
    a.st[i].s0 = m[i].s0;
    a.st[i].s1 = m[i].s1;
  }
  return a;
}

int check_mask(const xl_mask_stage* m, int nin, int batch = 1) {
  for (int i = 0; i < 2; ++i) {
    const int k = m[i].kind;
    if (k < XL_MASK_NONE || k > XL_MASK_AMP) return fail(XL_EINVAL, "unknown mask kind");
    const int pf = m[i].forms;
    if (pf & ~63) return fail(XL_EINVAL, "unknown mask parameter form bits");
    for (int j = 0; j < 2; ++j) {
      const int f = pf >> (3 * j) & 7;
      if (f && !(f & XL_PF_F64)) return fail(XL_EINVAL, "XL_PF_SCALAR / XL_PF_UNIT_ANGLE need XL_PF_F64");
    }
    if (pf && batch > 1) return fail(XL_EINVAL, "float64 mask parameter forms need batch <= 1");
    if (k == XL_MASK_NONE) continue;
    if ((k == XL_MASK_SLM || k == XL_MASK_SSLM || k == XL_MASK_LCD || k == XL_MASK_LP || k == XL_MASK_AMP) &&
        !m[i].m0)
      return fail(XL_EINVAL, "mask parameter m0 is NULL");
    if ((k == XL_MASK_SSLM || k == XL_MASK_LCD) && !m[i].m1) return fail(XL_EINVAL, "mask parameter m1 is NULL");
    if (k == XL_MASK_AMP && m[i].m1 && nin < 2) return fail(XL_EINVAL, "an Ey amplitude needs Ex and Ey");
    if (k != XL_MASK_SLM && k != XL_MASK_AMP && nin < 2) return fail(XL_EINVAL, "polarisation masks need Ex and Ey");
  }
  return XL_OK;
}

// ---------------------------------------------------------------- RS
int rs_fft_len(int n) { return std::max(64, next_pow2(2 * (int64_t)n - 1)); }

// complex64 elements of one problem's workspace block: tiles + kernel spectrum
int64_t rs_ws_elems(int n, int ncomp) {
  const int64_t P = rs_fft_len(n);
  const int64_t nr4 = ((n + 3) / 4) * 4;
  return ncomp * nr4 * P + (P / 2) * P + P;
}

size_t rs_ws(int n, int ncomp, int batch = 1) {
  return (size_t)std::max(1, batch) * (size_t)rs_ws_elems(n, ncomp) * sizeof(float2) + 256;
}

// Byte offsets in the spec buffer of spec_mode 1 / 2 (from its 256-aligned
// start): the ncomp P x P input spectra, two derivative spectra in the Hhat
// layout, two arrays of per-unit partial sums; [5] is the total.
struct RsSpecLayout {
  int64_t off[6];
  int units;   // padded (column, component) units of the column pass
};
RsSpecLayout rs_spec_layout(int n, int ncomp) {
  const int64_t P = rs_fft_len(n);
  const int tpf = (int)(P / 16), cpb = tpf >= 64 ? 1 : 64 / tpf;   // Units<P>::CPB
  RsSpecLayout l;
  l.units = (int)((P * ncomp + cpb - 1) / cpb * cpb);
  auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  const int64_t hh = ((P / 2) * P + P) * (int64_t)sizeof(float2);
  l.off[0] = 0;
  l.off[1] = up(ncomp * P * P * (int64_t)sizeof(float2));
  l.off[2] = l.off[1] + up(hh);
  l.off[3] = l.off[2] + up(hh);
  l.off[4] = l.off[3] + up((int64_t)l.units * 8);
  l.off[5] = l.off[4] + up((int64_t)l.units * 8);
  return l;
}

// ---------------------------------------------------------------- CZT
struct AxisPlan {
  int m, M, q;
  double d1_dm, dl_dm, th_turns;
};

int plan_axis(int m, int M, double f1, double f2, double Dm, bool allow3, AxisPlan* ap) {
  if (M < 2) return fail(XL_EINVAL, "CZT output length must be >= 2 (the reference mis-broadcasts M_out=1)");
  if (m < 2) return fail(XL_EINVAL, "CZT input length must be >= 2");
  const int64_t mp = (int64_t)m + M - 1;
  if ((mp & (mp - 1)) == 0)
    return fail(XL_EINVAL, "CZT with m + M_out - 1 = " + std::to_string(mp) +
                               " a power of two is rejected (the reference's window slice is one sample short)");
  // Bluestein length: the smaller of 2^k and (allow3) 3 * 2^k (768..3072)
  // covering mp; any length >= mp gives the same linear convolution
  int q = std::max(64, next_pow2(mp));
  // a power of two that holds both m and M in its halves runs the zero-half
  // passes and the staged column I/O, and then beats the shorter 3 * 2^k
  const bool half = 2 * std::max(m, M) <= q;
  for (int q3 = 768; allow3 && !half && q3 <= 3072; q3 *= 2)
    if (q3 >= mp && q3 < q) {
      q = q3;
      break;
    }
  if (q > xl::MAX_FFT)
    return fail(XL_EINVAL, "CZT needs FFT length " + std::to_string(q) + " > " + std::to_string(xl::MAX_FFT) +
                               " (m + M_out - 1 must be below 4096)");
  // wave_optics.py:435-437, :391-393
  const double D1 = f1 + (M * Dm + f2 - f1) / (2.0 * M);
  const double D2 = f2 + (M * Dm + f2 - f1) / (2.0 * M);
  const double delta = (D2 - D1) / M;
  double th = -2.0 * M_PI * (D2 - D1) / (M * Dm);
  th = std::atan2(std::sin(th), std::cos(th));   // principal branch of angle(W)
  ap->m = m;
  ap->M = M;
  ap->q = q;
  ap->d1_dm = D1 / Dm;
  ap->dl_dm = delta / Dm;
  ap->th_turns = th / (2.0 * M_PI);
  return XL_OK;
}

constexpr int kFtabBatch = 4;   // batch size from which the plane factors are tabulated

struct CztPlan {
  AxisPlan x, y;       // forward axes (x: nx_in -> mx, y: ny_in -> my)
  bool ftab;
  size_t ftab_bytes;
  int ncomp, batch;
  int rin, lin_x, lout_x, lout_y;
  size_t tables_bytes, tiles_bytes;
};

int plan_czt(const xl_czt_params* p, CztPlan* cp) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (p->ny_in < 2 || p->nx_in < 2 || p->my < 1 || p->mx < 1) return fail(XL_EINVAL, "bad CZT plane sizes");
  if (!(p->dm != 0.0) || !std::isfinite(p->dm)) return fail(XL_EINVAL, "Dm must be finite and non-zero");
  // 3 * 2^k lengths may serve the column pass (y axis) where the power of two
  // would not be half empty; the row pass keeps power-of-two lengths
  // (measured faster there: three passes instead of four)
  int rc = plan_axis(p->nx_in, p->mx, p->fx1, p->fx2, p->dm, false, &cp->x);
  if (rc) return rc;
  rc = plan_axis(p->ny_in, p->my, p->fy1, p->fy2, p->dm, true, &cp->y);
  if (rc) return rc;
  cp->ncomp = p->ncomp;
  if (!p->adjoint) {
    cp->rin = p->ny_in; cp->lin_x = p->nx_in; cp->lout_x = p->mx; cp->lout_y = p->my;
  } else {
    cp->rin = p->my; cp->lin_x = p->mx; cp->lout_x = p->nx_in; cp->lout_y = p->ny_in;
  }
  if (p->batch < 0 || p->batch > 65535) return fail(XL_EINVAL, "batch must be 0..65535");
  cp->batch = std::max(1, (int)p->batch);
  // batches of RS-mode CZTs share tables of the plane factors (F, F0)
  cp->ftab = cp->batch >= kFtabBatch && p->mode == XL_CZT_RS;
  cp->ftab_bytes = cp->ftab ? ((size_t)cp->rin * cp->lin_x + (size_t)cp->lout_y * cp->lout_x) * sizeof(float2) : 0;
  cp->tables_bytes = (size_t)2 * 3 * xl::MAX_FFT * sizeof(float2);
  cp->tiles_bytes = (size_t)cp->batch * cp->ncomp * ((cp->rin + 3) / 4 * 4) * ((cp->lout_x + 3) / 4 * 4) * sizeof(float2);
  return XL_OK;
}

template <typename F>
int dispatch_q(int q, F&& f) {
  switch (q) {
    case 64: f(std::integral_constant<int, 64>()); return XL_OK;
    case 128: f(std::integral_constant<int, 128>()); return XL_OK;
    case 256: f(std::integral_constant<int, 256>()); return XL_OK;
    case 512: f(std::integral_constant<int, 512>()); return XL_OK;
    case 1024: f(std::integral_constant<int, 1024>()); return XL_OK;
    case 2048: f(std::integral_constant<int, 2048>()); return XL_OK;
    case 4096: f(std::integral_constant<int, 4096>()); return XL_OK;
    default: return fail(XL_EINVAL, "unsupported FFT length " + std::to_string(q));
  }
}

// Bluestein lengths: 2^k (64..4096) and 3 * 2^k (768..3072).
template <typename F>
int dispatch_czt_q(int q, F&& f) {
  switch (q) {
    case 768: f(std::integral_constant<int, 768>()); return XL_OK;
    case 1536: f(std::integral_constant<int, 1536>()); return XL_OK;
    case 3072: f(std::integral_constant<int, 3072>()); return XL_OK;
    default: return dispatch_q(q, std::forward<F>(f));
  }
}

// One launch for both axis tables when the x axis has a power-of-two length
// (512..4096) and the y axis a 3 * 2^k length; false otherwise.
bool setup_pair_launch(int qx, int qy, const xl::CztSetupPair& pair, hipStream_t st) {
  auto with_y = [&](auto cx) {
    constexpr int QX = decltype(cx)::value;
    switch (qy) {
      case 768: xl::czt_setup2_launch<QX, 768>(pair, st); return true;
      case 1536: xl::czt_setup2_launch<QX, 1536>(pair, st); return true;
      case 3072: xl::czt_setup2_launch<QX, 3072>(pair, st); return true;
      default: return false;
    }
  };
  switch (qx) {
    case 512: return with_y(std::integral_constant<int, 512>());
    case 1024: return with_y(std::integral_constant<int, 1024>());
    case 2048: return with_y(std::integral_constant<int, 2048>());
    case 4096: return with_y(std::integral_constant<int, 4096>());
    default: return false;
  }
}

char* align256(char* p) { return (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255); }

}  // namespace

extern "C" {

const char* xl_last_error(void) { return g_err.c_str(); }
int xl_abi_version(void) { return 10; }
int xl_max_fft_length(void) { return xl::MAX_FFT; }
int xl_init(void) { return ensure_init(); }

size_t xl_rs_hhat_bytes(int32_t n) {
  if (n < 2) return 0;
  const int64_t P = rs_fft_len(n);
  return (size_t)((P / 2) * P + P) * sizeof(float2);
}

int64_t xl_rs_spec_partials(int32_t n, int32_t ncomp, int32_t i, int32_t* count) {
  if (!count || i < 0 || i > 1 || xl_rs_spec_bytes(n, ncomp) == 0) return -1;
  const RsSpecLayout l = rs_spec_layout(n, ncomp);
  *count = l.units;
  return l.off[3 + i];
}

size_t xl_rs_spec_bytes(int32_t n, int32_t ncomp) {
  if (n < 2 || ncomp < 1 || ncomp > 3 || rs_fft_len(n) > xl::MAX_FFT) return 0;
  return (size_t)rs_spec_layout(n, ncomp).off[5] + 256;
}

size_t xl_rs_workspace_bytes(int32_t n, int32_t ncomp) {
  if (n < 2 || ncomp < 1 || ncomp > 3) return 0;
  return rs_ws(n, ncomp);
}

size_t xl_rs_batch_workspace_bytes(int32_t n, int32_t ncomp, int32_t batch) {
  if (n < 2 || ncomp < 1 || ncomp > 3 || batch < 0 || batch > 65535) return 0;
  return rs_ws(n, ncomp, batch);
}

int xl_rs_propagate(const xl_rs_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!device_ready()) return fail(XL_EINVAL, "xl_init() has not run on this device");
  const int n = p->n;
  if (n < 2) return fail(XL_EINVAL, "grid side must be >= 2");
  const int P = rs_fft_len(n);
  if (P > xl::MAX_FFT) return fail(XL_EINVAL, "grid side " + std::to_string(n) + " needs FFT length > 4096");
  if (p->nin < 1 || p->nin > 3) return fail(XL_EINVAL, "nin must be 1..3");
  if (p->ez_mode != XL_EZ_NONE && p->ez_mode != XL_EZ_VRS) return fail(XL_EINVAL, "RS supports XL_EZ_NONE / XL_EZ_VRS");
  if (p->ez_mode == XL_EZ_VRS && p->nin < 2) return fail(XL_EINVAL, "Ez construction needs Ex and Ey");
  if (!p->in || !p->out || !p->x || !p->y) return fail(XL_EINVAL, "null pointer argument");
  if (int rc = check_mask(p->mask, p->nin, p->batch)) return rc;
  if (p->adjoint && (p->mask[0].kind || p->mask[1].kind || p->ez_mode))
    return fail(XL_EINVAL, "the adjoint convolution takes no masks / Ez construction (use xl_pre_backward)");
  if (p->hhat_mode < 0 || p->hhat_mode > 2) return fail(XL_EINVAL, "hhat_mode must be 0, 1 or 2");
  if (p->hhat_mode && (!p->hhat || p->kernel_dz))
    return fail(XL_EINVAL, "hhat_mode 1 / 2 needs the hhat buffer and kernel_dz = 0");
  const int ncomp = p->ez_mode == XL_EZ_VRS ? 3 : p->nin;
  if (p->batch < 0 || p->batch > 65535) return fail(XL_EINVAL, "batch must be 0..65535");
  if (p->spec_mode < 0 || p->spec_mode > 2) return fail(XL_EINVAL, "spec_mode must be 0, 1 or 2");
  if (p->spec_mode && (!p->spec || p->batch > 1 || p->kernel_dz))
    return fail(XL_EINVAL, "spec_mode 1 / 2 needs the spec buffer, one problem and kernel_dz = 0");
  if (p->spec_mode == 1 && (p->spec_grads < 0 || p->spec_grads > 3))
    return fail(XL_EINVAL, "spec_mode 1 takes spec_grads 0..3");
  if (p->spec_mode == 2 && ((p->spec_grads & 3) == 0 || p->spec_grads < 0 || p->spec_grads > 15 ||
                             (!p->grad_out && !(p->spec_grads & XL_RS_SPEC_PARTIALS))))
    return fail(XL_EINVAL, "spec_mode 2 needs spec_grads 1..3 (| XL_RS_SPEC_PREBUILT | XL_RS_SPEC_PARTIALS) and "
                           "grad_out unless XL_RS_SPEC_PARTIALS");
  if (p->spec_dg && (!p->spec_mode || (p->spec_mode == 2 && !(p->spec_grads & XL_RS_SPEC_PREBUILT))))
    return fail(XL_EINVAL, "spec_dg needs spec_mode 1, or spec_mode 2 with XL_RS_SPEC_PREBUILT");
  const int batch = std::max(1, (int)p->batch);
  const int64_t n2 = (int64_t)n * n;
  if (batch > 1) {
    if (p->in_batch_stride < 0 || p->mask_batch_stride < 0)
      return fail(XL_EINVAL, "input / mask batch strides must be >= 0 (0 shares one array)");
    if (p->mask_batch_stride > 0 && p->mask_batch_stride < n2)
      return fail(XL_EINVAL, "per-member mask parameters must not overlap (mask_batch_stride >= n^2)");
    // members write their outputs concurrently: overlapping output blocks would race
    if (p->out_batch_stride < (int64_t)ncomp * n2)
      return fail(XL_EINVAL, "out_batch_stride must be >= ncomp n^2 (members write disjoint outputs)");
    if (p->intensity && p->intensity_batch_stride < n2)
      return fail(XL_EINVAL, "intensity_batch_stride must be >= n^2 (members write disjoint intensities)");
  }
  const size_t need = rs_ws(n, ncomp, batch);
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "RS workspace too small: need " + std::to_string(need));
  char* w = align256((char*)ws);
  const int64_t nr4 = ((n + 3) / 4) * 4;
  xl::RsArgs a;
  a.in = (const float2*)p->in;
  a.in_cs = p->in_comp_stride;
  a.out = (float2*)p->out;
  a.out_cs = p->out_comp_stride;
  a.intensity = p->intensity;
  a.t = (float2*)w;
  a.hr = p->hhat_mode ? (float2*)p->hhat : a.t + ncomp * nr4 * P;
  a.hmid = a.hr + (int64_t)(P / 2) * P;
  a.hstored = p->hhat_mode ? 1 : 0;
  a.hreuse = p->hhat_mode == 2 ? 1 : 0;
  a.x = p->x;
  a.y = p->y;
  a.z = p->z;
  a.k = p->k;
  a.z_dev = p->z_dev;
  if (p->kernel_dz < 0 || p->kernel_dz > 2) return fail(XL_EINVAL, "kernel_dz must be 0, 1 (dH/dz) or 2 (dH/dk)");
  a.kernel_dz = p->kernel_dz;
  a.n = n;
  a.nin = p->nin;
  a.ncomp = ncomp;
  a.ez_mode = p->ez_mode;
  a.adjoint = p->adjoint;
  a.mask = to_mask(p->mask);
  a.scale = (float)(p->dx * p->dy / ((double)P * (double)P));
  a.batch = batch;
  a.in_bs = batch > 1 ? p->in_batch_stride : 0;
  a.out_bs = batch > 1 ? p->out_batch_stride : 0;
  a.inten_bs = batch > 1 ? p->intensity_batch_stride : 0;
  a.mask_bs = batch > 1 ? p->mask_batch_stride : 0;
  a.t_bs = rs_ws_elems(n, ncomp);
  a.hr_bs = p->hhat_mode ? (int64_t)(P / 2) * P + P : a.t_bs;   // hhat holds one spectrum per member
  a.spec_mode = p->spec_mode;
  a.spec = nullptr;
  a.dg[0] = a.dg[1] = nullptr;
  a.dg_built = p->spec_mode == 2 && (p->spec_grads & XL_RS_SPEC_PREBUILT) ? 1 : 0;
  a.partials_only = p->spec_mode == 2 && (p->spec_grads & XL_RS_SPEC_PARTIALS) ? 1 : 0;
  a.gpart[0] = a.gpart[1] = nullptr;
  a.grad_out = p->grad_out;
  a.qf_out = p->quality_factor;
  a.qf_rmax = p->qf_rmax;
  a.qf_dr = p->qf_dr_real;
  a.qf_wl = p->qf_wavelength;
  if (p->quality_factor && !(p->qf_dr_real > 0.0)) return fail(XL_EINVAL, "the quality factor needs qf_dr_real > 0");
  if (p->spec_mode) {
    const RsSpecLayout l = rs_spec_layout(n, ncomp);
    char* sb = align256((char*)p->spec);
    char* db = p->spec_dg ? align256((char*)p->spec_dg) : sb;   // where the derivative spectra live
    a.spec = (float2*)sb;
    for (int i = 0; i < 2; ++i)
      if (p->spec_grads >> i & 1) {
        // a forward call with spec_dg builds nothing: the other call did
        if (!(p->spec_mode == 1 && p->spec_dg)) a.dg[i] = (float2*)(db + l.off[1 + i]);
        if (p->spec_mode == 2) a.gpart[i] = (double*)(sb + l.off[3 + i]);
      }
  }
  hipStream_t st = (hipStream_t)stream;
  int rc = dispatch_q(P, [&](auto c) { (void)xl::rs_launch<decltype(c)::value>(a, st); });
  if (rc) return rc;
  return hip_check(hipGetLastError(), "RS launch");
}

int xl_rs_propagate_noalloc(const xl_rs_params* p, void* stream) {
  static_assert(XL_RS_NOALLOC_MAX_N == 1024, "the scratch size below is the n = 1024 workspace");
  static_assert(XL_RS_NOALLOC_SCRATCH_BYTES >= (3 * 1024 * 2048 + 1024 * 2048 + 2048) * 8 + 256,
                "scratch smaller than the n = 1024, 3-component workspace");
  if (!p) return fail(XL_EINVAL, "null params");
  if (p->n > XL_RS_NOALLOC_MAX_N)
    return fail(XL_EINVAL, "xl_rs_propagate_noalloc takes grids up to " + std::to_string(XL_RS_NOALLOC_MAX_N) +
                               " per side (its scratch holds " + std::to_string(XL_RS_NOALLOC_SCRATCH_BYTES) +
                               " bytes); use xl_rs_propagate with a workspace for n = " + std::to_string(p->n));
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail(XL_EHIP, "hipGetDevice failed");
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (!g_init_done[dev]) return fail(XL_EINVAL, "xl_init() has not run on this device");
    ws = g_scratch[dev];
  }
  const int ncomp = p->ez_mode == XL_EZ_VRS ? 3 : std::max(1, (int)p->nin);
  if (p->n >= 2 && ncomp <= 3 && rs_ws(p->n, ncomp, std::max(1, (int)p->batch)) > XL_RS_NOALLOC_SCRATCH_BYTES)
    return fail(XL_EINVAL, "this batch of " + std::to_string(p->n) + "^2 problems needs more than the " +
                               std::to_string(XL_RS_NOALLOC_SCRATCH_BYTES) + "-byte scratch of xl_rs_propagate_noalloc");
  // every kernel writes the scratch before reading it (zero-padded rows and
  // columns included), so nothing in it has to start at zero
  return xl_rs_propagate(p, ws, XL_RS_NOALLOC_SCRATCH_BYTES, stream);
}

size_t xl_czt_workspace_bytes(const xl_czt_params* p) {
  CztPlan cp;
  if (plan_czt(p, &cp)) return 0;
  return cp.tables_bytes + cp.tiles_bytes + cp.ftab_bytes + 768;
}

static dim3 grid_for(int64_t npix, int batch = 1);

int xl_czt(const xl_czt_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!device_ready()) return fail(XL_EINVAL, "xl_init() has not run on this device");
  CztPlan cp;
  if (int rc = plan_czt(p, &cp)) return rc;
  if (!p->in || !p->out || !p->x || !p->y || !p->xout || !p->yout) return fail(XL_EINVAL, "null pointer argument");
  if (p->nin < 1 || p->nin > 3 || p->ncomp < 1 || p->ncomp > 3) return fail(XL_EINVAL, "nin / ncomp must be 1..3");
  if (p->mode < XL_CZT_RS || p->mode > XL_CZT_PLAIN) return fail(XL_EINVAL, "unknown CZT mode");
  if (p->mode == XL_CZT_HNA && !p->adjoint && (p->nin < 2 || p->ncomp != 3))
    return fail(XL_EINVAL, "high-NA CZT takes Ex, Ey and returns 3 components");
  if (p->ez_mode != XL_EZ_NONE && p->ez_mode != XL_EZ_VCZT) return fail(XL_EINVAL, "CZT supports XL_EZ_NONE / XL_EZ_VCZT");
  if (!p->adjoint && p->ez_mode == XL_EZ_VCZT && (p->nin < 2 || p->ncomp != 3))
    return fail(XL_EINVAL, "Ez construction needs Ex, Ey and 3 output components");
  if (!p->adjoint && p->ez_mode == XL_EZ_NONE && p->mode != XL_CZT_HNA && p->ncomp != p->nin)
    return fail(XL_EINVAL, "ncomp must equal nin without Ez construction");
  if (p->adjoint && (p->mask[0].kind || p->mask[1].kind || p->ez_mode || p->nin != p->ncomp))
    return fail(XL_EINVAL, "the adjoint CZT takes ncomp cotangents and no masks / Ez (use xl_pre_backward)");
  if (int rc = check_mask(p->mask, p->nin, p->batch)) return rc;
  const size_t need = cp.tables_bytes + cp.tiles_bytes + cp.ftab_bytes + 768;
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "CZT workspace too small: need " + std::to_string(need));
  char* w = align256((char*)ws);
  float2* tab = (float2*)w;
  float2* tiles = (float2*)align256((char*)(tab + 6 * xl::MAX_FFT));
  hipStream_t st = (hipStream_t)stream;

  // axis tables: x in tab[0..3Q), y in tab[3Q..6Q)
  xl::CztArgs a;
  std::memset(&a, 0, sizeof(a));
  const AxisPlan* axp[2] = {&cp.x, &cp.y};
  xl::CztAxis* axo[2] = {&a.ax, &a.ay};
  xl::CztSetupPair pair;
  for (int i = 0; i < 2; ++i) {
    const AxisPlan& ap = *axp[i];
    xl::CztSetupArgs& s = pair.ax[i];
    s.pre = tab + i * 3 * xl::MAX_FFT;
    s.gf = s.pre + xl::MAX_FFT;
    s.post = s.gf + xl::MAX_FFT;
    s.m = ap.m;
    s.M = ap.M;
    s.q = ap.q;
    s.adjoint = p->adjoint;
    s.d1_dm = ap.d1_dm;
    s.dl_dm = ap.dl_dm;
    s.th_turns = ap.th_turns;
    xl::CztAxis& ax = *axo[i];
    ax.q = ap.q;
    ax.pre = s.pre;
    ax.gf = s.gf;
    ax.post = s.post;
    ax.lin = p->adjoint ? ap.M : ap.m;
    ax.lout = p->adjoint ? ap.m : ap.M;
    // zero-half passes: power-of-two length with the input and the output in its lower half
    ax.zl = (ap.q & (ap.q - 1)) == 0 && 2 * std::max(ap.m, ap.M) <= ap.q;
  }
  if (p->tables_mode < 0 || p->tables_mode > 2) return fail(XL_EINVAL, "tables_mode must be 0, 1 or 2");
  if (p->tables_mode != 2) {
    if (cp.x.q == cp.y.q) {
      if (int rc = dispatch_czt_q(cp.x.q, [&](auto c) { xl::czt_setup_launch<decltype(c)::value>(pair, 2, st); }))
        return rc;
    } else if (setup_pair_launch(cp.x.q, cp.y.q, pair, st)) {
      // both axes in one launch
    } else {
      for (int i = 0; i < 2; ++i) {
        xl::CztSetupPair one;
        one.ax[0] = pair.ax[i];
        if (int rc = dispatch_czt_q(axp[i]->q, [&](auto c) { xl::czt_setup_launch<decltype(c)::value>(one, 1, st); }))
          return rc;
      }
    }
    if (p->tables_mode == 1) return hip_check(hipGetLastError(), "CZT table staging");
  }
  a.in = (const float2*)p->in;
  a.in_cs = p->in_comp_stride;
  a.nin = p->nin;
  a.out = (float2*)p->out;
  a.out_cs = p->out_comp_stride;
  a.t = tiles;
  a.rin = cp.rin;
  a.ncomp = p->ncomp;
  a.ez_mode = p->adjoint ? XL_EZ_NONE : p->ez_mode;
  a.z = p->z;
  a.k = p->k;
  a.hna_radius = p->hna_radius;
  a.hna_f = p->hna_f;
  a.hna_stm2 = p->hna_radius * p->hna_radius / (p->hna_radius * p->hna_radius + p->hna_f * p->hna_f);
  a.mask = to_mask(p->mask);
  const float2 cte = make_float2((float)p->cte_re, (float)p->cte_im);
  const float2 ctec = make_float2((float)p->cte_re, (float)-p->cte_im);
  if (!p->adjoint) {
    a.px = p->x; a.py = p->y; a.qx = p->xout; a.qy = p->yout;
    a.cte = cte;
    if (p->mode == XL_CZT_RS) { a.pre_mode = xl::PRE_RS; a.post_mode = xl::POST_RS; }
    else if (p->mode == XL_CZT_HNA) { a.pre_mode = xl::PRE_HNA; a.post_mode = xl::POST_CONST; }
    else { a.pre_mode = xl::PRE_NONE; a.post_mode = xl::POST_NONE; }
  } else {
    a.px = p->xout; a.py = p->yout; a.qx = p->x; a.qy = p->y;
    a.cte = ctec;
    if (p->mode == XL_CZT_RS) { a.pre_mode = xl::PRE_CONJ_RS; a.post_mode = xl::POST_CONJ_RS; }
    else if (p->mode == XL_CZT_HNA) { a.pre_mode = xl::PRE_CONST; a.post_mode = xl::POST_NONE; }
    else { a.pre_mode = xl::PRE_NONE; a.post_mode = xl::POST_NONE; }
  }
  const int batch = std::max(1, (int)p->batch);
  if (cp.ftab) {
    float2* ft = (float2*)align256((char*)(tiles + cp.tiles_bytes / sizeof(float2)));
    float2* f0 = ft + (size_t)cp.rin * cp.lin_x;
    xl::KTimer t(xl::K_CZT_SETUP, st);
    hipLaunchKernelGGL(xl::czt_ftab_kernel, grid_for(cp.ftab_bytes / sizeof(float2)), dim3(256), 0, st, a, ft, f0);
    a.ftab = ft;
    a.f0tab = f0;
  }
  a.in_bs = batch > 1 ? p->in_batch_stride : 0;
  a.out_bs = batch > 1 ? p->out_batch_stride : 0;
  a.mask_bs = batch > 1 ? p->mask_batch_stride : 0;
  a.rin4 = (cp.rin + 3) / 4 * 4;
  a.t_cs = (int64_t)a.rin4 * ((cp.lout_x + 3) / 4 * 4);
  a.t_bs = (int64_t)cp.ncomp * a.t_cs;
  if (int rc = dispatch_czt_q(a.ax.q, [&](auto c) { xl::czt_rows_launch<decltype(c)::value>(a, batch, st); }))
    return rc;
  if (int rc = dispatch_czt_q(a.ay.q, [&](auto c) { xl::czt_cols_launch<decltype(c)::value>(a, batch, st); }))
    return rc;
  return hip_check(hipGetLastError(), "CZT launch");
}


// ---------------------------------------------------------------- zoom-GEMM CZT
namespace {
struct GemmPlan {
  CztPlan cz;
  int nxp, nyp, mxp, myp;            // padded (x32) lengths
  int a1_rows, a1_ld, a2_rows, a2_ld;  // kernel matrices of stage 1 / 2
  int b_rowsp, b_ld;                 // pre-op planes: rows per component, row stride
  int64_t c1_ld;                     // stage-1 output row stride
  int c1_rows;
  int wn[2], splits[2];              // tile width (32 * wn) and split-K factor per stage
  size_t ka_bytes, kb_bytes, g_bytes, c1_bytes, part_bytes;
};

int r32(int v) { return (v + 31) / 32 * 32; }

// Tile width and split-K factor for an M x N x K product: the grid should
// fill the 256 CUs (blocks of 4 x WN waves at 4 waves per SIMD: WN = 2 fits
// two blocks per CU, WN = 3 or 4 one).  Split-K costs an fp32 partial pass
// that grows with the split count: a split is charged 0.1 + 0.01 per split
// against the filled fraction, scaled down when the partials are small
// (under 32 MiB).
void choose_gemm_shape(int M, int N, int K, int* wn_out, int* splits_out) {
  constexpr int kCUs = 256;
  const int tm = (M + xl::GBM - 1) / xl::GBM;
  const int ksteps = K / xl::GBK;
  const int smax = std::max(1, std::min(16, ksteps / 4));
  double best = -1.0;
  for (int wn = 4; wn >= 2; --wn) {
    const int tn = (N + 32 * wn - 1) / (32 * wn);
    const int slots = kCUs * (wn == 2 ? 2 : 1);
    for (int sp = 1; sp <= smax; ++sp) {
      const int64_t blocks = (int64_t)tm * tn * sp;
      const int64_t rounds = (blocks + slots - 1) / slots;
      const double part = (double)sp * M * N * sizeof(float2);
      const double charge = sp > 1 ? (0.1 + 0.01 * sp) * std::min(1.0, part / (32.0 * 1048576.0)) : 0.0;
      const double score = (double)blocks / (double)(rounds * slots) - charge;
      if (score > best + 1e-9) {
        best = score;
        *wn_out = wn;
        *splits_out = sp;
      }
    }
  }
}

int plan_gemm(const xl_czt_params* p, GemmPlan* g) {
  if (int rc = plan_czt(p, &g->cz)) return rc;
  g->nxp = r32(p->nx_in); g->nyp = r32(p->ny_in); g->mxp = r32(p->mx); g->myp = r32(p->my);
  const int nc = p->ncomp;
  if (!p->adjoint) {
    g->a1_rows = p->mx; g->a1_ld = g->nxp;        // K_x
    g->a2_rows = p->my; g->a2_ld = g->nyp;        // K_y
    g->b_rowsp = g->nyp; g->b_ld = g->nxp;        // G
    g->c1_rows = p->mx; g->c1_ld = (int64_t)nc * g->nyp;
  } else {
    g->a1_rows = p->nx_in; g->a1_ld = g->mxp;     // K_x^H
    g->a2_rows = p->ny_in; g->a2_ld = g->myp;     // K_y^H
    g->b_rowsp = g->myp; g->b_ld = g->mxp;        // G'
    g->c1_rows = p->nx_in; g->c1_ld = (int64_t)nc * g->myp;
  }
  g->ka_bytes = (size_t)4 * g->a1_rows * g->a1_ld * 2;
  g->kb_bytes = (size_t)4 * g->a2_rows * g->a2_ld * 2;
  g->g_bytes = (size_t)4 * nc * g->b_rowsp * g->b_ld * 2;
  g->c1_bytes = (size_t)4 * g->c1_rows * g->c1_ld * 2;
  // stage 1: M = a1_rows, N = nc * b_rowsp, K = b_ld; stage 2: M = a2_rows, N = nc * a1_rows, K = b_rowsp
  const int sm[2] = {g->a1_rows, g->a2_rows}, sn[2] = {nc * g->b_rowsp, nc * g->a1_rows}, sk[2] = {g->b_ld, g->b_rowsp};
  g->part_bytes = 0;
  for (int i = 0; i < 2; ++i) {
    choose_gemm_shape(sm[i], sn[i], sk[i], &g->wn[i], &g->splits[i]);
    if (g->splits[i] > 1)
      g->part_bytes = std::max(g->part_bytes, (size_t)g->splits[i] * sm[i] * sn[i] * sizeof(float2));
  }
  return XL_OK;
}

size_t gemm_ws(const GemmPlan& g) {
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  return al(g.ka_bytes) + al(g.kb_bytes) + al(g.g_bytes) + al(g.c1_bytes) + al(g.part_bytes) + 256;
}

// Launches C = A B^T with the stage's tile width and split-K factor (then the
// partial reduction when split).
void launch_cgemm(xl::GemmArgs& a, int wn, int splits, hipStream_t st) {
  const dim3 grid((a.N + 32 * wn - 1) / (32 * wn), (a.M + xl::GBM - 1) / xl::GBM, splits);
  if (wn == 4) hipLaunchKernelGGL(xl::cgemm_kernel<4>, grid, dim3(1024), 0, st, a);
  else if (wn == 3) hipLaunchKernelGGL(xl::cgemm_kernel<3>, grid, dim3(768), 0, st, a);
  else hipLaunchKernelGGL(xl::cgemm_kernel<2>, grid, dim3(512), 0, st, a);
  if (splits > 1)
    hipLaunchKernelGGL(xl::cgemm_reduce_kernel, grid_for((int64_t)a.M * a.N), dim3(256), 0, st, a, splits);
}

xl::Planes planes_at(char* base, int rows, int64_t ld) {
  xl::Planes pl;
  pl.p = (uint16_t*)base;
  pl.ld = ld;
  pl.ps = (int64_t)rows * ld;
  return pl;
}
}  // namespace

size_t xl_czt_gemm_workspace_bytes(const xl_czt_params* p) {
  GemmPlan g;
  if (plan_gemm(p, &g)) return 0;
  return gemm_ws(g);
}

int xl_czt_gemm(const xl_czt_params* p, void* ws, size_t ws_bytes, void* stream) {
  GemmPlan g;
  if (int rc = plan_gemm(p, &g)) return rc;
  if (p->batch > 1) return fail(XL_EINVAL, "batched CZT runs in the Bluestein form (xl_czt)");
  if (p->tables_mode != 0) return fail(XL_EINVAL, "tables_mode is a Bluestein-form option (xl_czt); it must be 0 here");
  if (!p->in || !p->out || !p->x || !p->y || !p->xout || !p->yout) return fail(XL_EINVAL, "null pointer argument");
  if (p->nin < 1 || p->nin > 3 || p->ncomp < 1 || p->ncomp > 3) return fail(XL_EINVAL, "nin / ncomp must be 1..3");
  if (p->mode < XL_CZT_RS || p->mode > XL_CZT_PLAIN) return fail(XL_EINVAL, "unknown CZT mode");
  if (p->mode == XL_CZT_HNA && !p->adjoint && (p->nin < 2 || p->ncomp != 3))
    return fail(XL_EINVAL, "high-NA CZT takes Ex, Ey and returns 3 components");
  if (p->ez_mode != XL_EZ_NONE && p->ez_mode != XL_EZ_VCZT) return fail(XL_EINVAL, "CZT supports XL_EZ_NONE / XL_EZ_VCZT");
  if (!p->adjoint && p->ez_mode == XL_EZ_VCZT && (p->nin < 2 || p->ncomp != 3))
    return fail(XL_EINVAL, "Ez construction needs Ex, Ey and 3 output components");
  if (!p->adjoint && p->ez_mode == XL_EZ_NONE && p->mode != XL_CZT_HNA && p->ncomp != p->nin)
    return fail(XL_EINVAL, "ncomp must equal nin without Ez construction");
  if (p->adjoint && (p->mask[0].kind || p->mask[1].kind || p->ez_mode || p->nin != p->ncomp))
    return fail(XL_EINVAL, "the adjoint CZT takes ncomp cotangents and no masks / Ez (use xl_pre_backward)");
  if (int rc = check_mask(p->mask, p->nin, p->batch)) return rc;
  const size_t need = gemm_ws(g);
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "zoom-GEMM workspace too small: need " + std::to_string(need));
  if (!device_ready()) return fail(XL_EINVAL, "xl_init() has not run on this device");
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* w = align256((char*)ws);
  xl::Planes ka = planes_at(w, g.a1_rows, g.a1_ld);
  w += al(g.ka_bytes);
  xl::Planes kb = planes_at(w, g.a2_rows, g.a2_ld);
  w += al(g.kb_bytes);
  xl::Planes gp = planes_at(w, p->ncomp * g.b_rowsp, g.b_ld);
  w += al(g.g_bytes);
  xl::Planes c1 = planes_at(w, g.c1_rows, g.c1_ld);
  w += al(g.c1_bytes);
  float2* part = g.part_bytes ? (float2*)w : nullptr;
  hipStream_t st = (hipStream_t)stream;
  const int nc = p->ncomp;

  // kernel matrices (x axis for stage 1, y axis for stage 2) and the pre-op
  // planes, prepared by one launch
  const AxisPlan* axp[2] = {&g.cz.x, &g.cz.y};
  xl::Planes* kd[2] = {&ka, &kb};
  const int krows[2] = {g.a1_rows, g.a2_rows};
  xl::KgenArgs k[2];
  int nbk[2];
  for (int i = 0; i < 2; ++i) {
    k[i].dst = *kd[i];
    k[i].m = axp[i]->m;
    k[i].M = axp[i]->M;
    k[i].rows = krows[i];
    k[i].cols = (int)kd[i]->ld;
    k[i].valid_cols = p->adjoint ? axp[i]->M : axp[i]->m;
    k[i].adjoint = p->adjoint;
    k[i].d1_dm = axp[i]->d1_dm;
    k[i].dl_dm = axp[i]->dl_dm;
    k[i].th_turns = axp[i]->th_turns;
    nbk[i] = (int)grid_for((int64_t)k[i].rows * k[i].cols / 2).x;
  }
  {
    xl::GemmPreArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = (const float2*)p->in;
    a.in_cs = p->in_comp_stride;
    a.nin = p->nin;
    a.ncomp = nc;
    a.rows = p->adjoint ? p->my : p->ny_in;
    a.cols = p->adjoint ? p->mx : p->nx_in;
    a.rowsp = g.b_rowsp;
    a.dst = gp;
    a.dst.ps = (int64_t)nc * g.b_rowsp * g.b_ld;
    a.ez_mode = p->adjoint ? XL_EZ_NONE : p->ez_mode;
    a.z = p->z;
    a.k = p->k;
    a.hna_radius = p->hna_radius;
    a.hna_f = p->hna_f;
    a.hna_stm2 = p->hna_radius * p->hna_radius / (p->hna_radius * p->hna_radius + p->hna_f * p->hna_f);
    a.mask = to_mask(p->mask);
    const float2 ctec = make_float2((float)p->cte_re, (float)-p->cte_im);
    if (!p->adjoint) {
      a.px = p->x; a.py = p->y;
      a.pre_mode = p->mode == XL_CZT_RS ? xl::PRE_RS : (p->mode == XL_CZT_HNA ? xl::PRE_HNA : xl::PRE_NONE);
    } else {
      a.px = p->xout; a.py = p->yout;
      a.cte = ctec;
      a.pre_mode = p->mode == XL_CZT_RS ? xl::PRE_CONJ_RS : (p->mode == XL_CZT_HNA ? xl::PRE_CONST : xl::PRE_NONE);
    }
    const int nbp = (int)grid_for((int64_t)g.b_rowsp * g.b_ld / 2).x;
    xl::KTimer t(xl::K_GEMM_PRE, st);
    hipLaunchKernelGGL(xl::gemm_prep_kernel, dim3(nbk[0] + nbk[1] + nbp), dim3(256), 0, st, k[0], k[1], a, nbk[0],
                       nbk[1]);
  }
  // stage 1: C1 = K1 G^T
  {
    xl::GemmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.A = {ka.p, ka.ps, ka.ld, g.a1_rows, 1 << 30, 0};
    a.B = {gp.p, (int64_t)nc * g.b_rowsp * g.b_ld, gp.ld, nc * g.b_rowsp, 1 << 30, 0};
    a.M = g.a1_rows;
    a.N = nc * g.b_rowsp;
    a.K = (int)g.b_ld;
    a.mode = 0;
    a.c = c1;
    a.partial = part;
    xl::KTimer t(xl::K_GEMM_X, st);
    launch_cgemm(a, g.wn[0], g.splits[0], st);
  }
  // stage 2: out = K2 B2^T, B2 row (c, j) = C1[j][c * rowsp ...]
  {
    xl::GemmArgs a;
    std::memset(&a, 0, sizeof(a));
    const int nb = g.a1_rows;   // stage-1 rows = output columns per component
    a.A = {kb.p, kb.ps, kb.ld, g.a2_rows, 1 << 30, 0};
    a.B = {c1.p, c1.ps, c1.ld, nc * nb, nb, (int64_t)g.b_rowsp};
    a.M = g.a2_rows;
    a.N = nc * nb;
    a.K = g.b_rowsp;
    a.mode = 1;
    a.out = (float2*)p->out;
    a.out_cs = p->out_comp_stride;
    a.nb = nb;
    a.z = p->z;
    a.k = p->k;
    if (!p->adjoint) {
      a.qx = p->xout; a.qy = p->yout;
      a.cte = make_float2((float)p->cte_re, (float)p->cte_im);
      a.post_mode = p->mode == XL_CZT_RS ? xl::POST_RS : (p->mode == XL_CZT_HNA ? xl::POST_CONST : xl::POST_NONE);
    } else {
      a.qx = p->x; a.qy = p->y;
      a.post_mode = p->mode == XL_CZT_RS ? xl::POST_CONJ_RS : xl::POST_NONE;
    }
    a.partial = part;
    xl::KTimer t(xl::K_GEMM_Y, st);
    launch_cgemm(a, g.wn[1], g.splits[1], st);
  }
  return hip_check(hipGetLastError(), "zoom-GEMM launch");
}

static int pre_common(const xl_pre_params* p, xl::PreArgs* a) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->in || !p->x || !p->y) return fail(XL_EINVAL, "null pointer argument");
  if (p->rows < 1 || p->cols < 1 || p->nin < 1 || p->nin > 3 || p->ncomp < 1 || p->ncomp > 3)
    return fail(XL_EINVAL, "bad pre-op shape");
  if (p->ez_mode < XL_EZ_NONE || p->ez_mode > XL_EZ_HNA) return fail(XL_EINVAL, "unknown Ez mode");
  if ((p->ez_mode != XL_EZ_NONE || p->hna) && (p->nin < 2 || p->ncomp != 3))
    return fail(XL_EINVAL, "Ez construction / lens map need Ex, Ey and 3 output components");
  if (int rc = check_mask(p->mask, p->nin, p->batch)) return rc;
  std::memset(a, 0, sizeof(*a));
  a->in = (const float2*)p->in;
  a->in_cs = p->in_comp_stride;
  a->nin = p->nin;
  a->ncomp = p->ncomp;
  a->gout = (const float2*)p->gout;
  a->out = (float2*)p->out;
  a->cs = p->comp_stride;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) a->gparam[i][j] = p->gparam[i][j];
  a->rows = p->rows;
  a->cols = p->cols;
  a->ez_mode = p->ez_mode;
  a->hna = p->hna;
  a->x = p->x;
  a->y = p->y;
  a->z = p->z;
  a->z_dev = p->z_dev;
  a->hna_radius = p->hna_radius;
  a->hna_f = p->hna_f;
  a->hna_stm2 = p->hna_radius * p->hna_radius / (p->hna_radius * p->hna_radius + p->hna_f * p->hna_f);
  if (p->out_planes < 0 || p->out_planes > 3) return fail(XL_EINVAL, "out_planes must be 0..3");
  a->out_planes = p->out_planes;
  a->mask = to_mask(p->mask);
  if (p->batch < 0 || p->batch > 65535) return fail(XL_EINVAL, "batch must be 0..65535");
  if (p->batch > 1) {
    a->in_bs = p->in_batch_stride;
    a->gout_bs = p->gout_batch_stride;
    a->out_bs = p->out_batch_stride;
    a->mask_bs = p->mask_batch_stride;
    a->gparam_bs = p->gparam_batch_stride;
  }
  return XL_OK;
}

static dim3 grid_for(int64_t npix, int batch) {
  int64_t b = (npix + 255) / 256;
  if (b > 65536) b = 65536;
  return dim3((unsigned)std::max<int64_t>(b, 1), (unsigned)std::max(batch, 1));
}

int xl_pre_forward(const xl_pre_params* p, void* stream) {
  xl::PreArgs a;
  if (int rc = pre_common(p, &a)) return rc;
  if (!p->out) return fail(XL_EINVAL, "null output");
  xl::KTimer t(xl::K_PRE_FWD, (hipStream_t)stream);
  hipLaunchKernelGGL(xl::pre_forward_kernel, grid_for((int64_t)p->rows * p->cols, p->batch), dim3(256), 0,
                     (hipStream_t)stream, a);
  return hip_check(hipGetLastError(), "pre_forward launch");
}

size_t xl_pre_scalar_grad_elems(int32_t rows, int32_t cols) {
  if (rows < 1 || cols < 1) return 0;
  return 1 + (size_t)grid_for((int64_t)rows * cols).x;
}

int xl_pre_backward(const xl_pre_params* p, void* stream) {
  xl::PreArgs a;
  if (int rc = pre_common(p, &a)) return rc;
  if (!p->gout) return fail(XL_EINVAL, "null cotangent");
  double* sbuf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      if (p->gparam[i][j] && (p->mask[i].forms >> (3 * j) & XL_PF_SCALAR)) {
        a.scal |= 1 << (2 * i + j);
        sbuf[2 * i + j] = (double*)p->gparam[i][j];
      }
  if (p->ez_dz) {
    if (p->ez_mode != XL_EZ_VRS || p->hna || p->batch > 1)
      return fail(XL_EINVAL, "ez_dz is the XL_EZ_VRS construction's distance term of one problem");
    a.scal |= 16;
    a.ezdz = sbuf[4] = p->ez_dz;
  }
  const bool any_gparam = p->gparam[0][0] || p->gparam[0][1] || p->gparam[1][0] || p->gparam[1][1];
  if (p->batch > 1 && any_gparam && p->gparam_batch_stride < (int64_t)p->rows * p->cols)
    return fail(XL_EINVAL, "batched mask-parameter gradients need one array per member (gparam_batch_stride >= "
                           "rows * cols): members writing one array would race");
  xl::KTimer t(xl::K_PRE_BWD, (hipStream_t)stream);
  const int64_t npix = (int64_t)p->rows * p->cols;
  const bool lone_slm = !p->hna && p->ez_mode == XL_EZ_NONE && p->nin == 1 && p->ncomp == 1 &&
                        p->mask[0].kind == XL_MASK_SLM && p->mask[1].kind == XL_MASK_NONE && !p->gparam[0][1] &&
                        p->mask[0].forms == 0 && p->out_planes <= 1;
  if (lone_slm) {
    hipLaunchKernelGGL(xl::pre_slm_backward_kernel, grid_for((npix + xl::kSlmUnroll - 1) / xl::kSlmUnroll, p->batch),
                       dim3(256), 0, (hipStream_t)stream, a);
  } else {
    const dim3 grid = grid_for(npix, p->batch);
    if (p->hna)
      hipLaunchKernelGGL(xl::pre_backward_kernel<2>, grid, dim3(256), 0, (hipStream_t)stream, a);
    else if (p->ez_mode != XL_EZ_NONE)
      hipLaunchKernelGGL(xl::pre_backward_kernel<1>, grid, dim3(256), 0, (hipStream_t)stream, a);
    else
      hipLaunchKernelGGL(xl::pre_backward_kernel<0>, grid, dim3(256), 0, (hipStream_t)stream, a);
    if (a.scal)
      hipLaunchKernelGGL(xl::pgrad_reduce_kernel, dim3(5), dim3(1024), 0, (hipStream_t)stream, sbuf[0], sbuf[1],
                         sbuf[2], sbuf[3], sbuf[4], p->ez_dz ? p->ez_dz_base : nullptr, (int)grid.x,
                         std::max(1, (int)p->ez_dz_base_n));
  }
  return hip_check(hipGetLastError(), "pre_backward launch");
}

int xl_slm_backward_adam(const xl_pre_params* p, const xl_adam_params* ad, void* stream) {
  xl::PreArgs a;
  if (int rc = pre_common(p, &a)) return rc;
  if (!p->gout) return fail(XL_EINVAL, "null cotangent");
  if (!ad) return fail(XL_EINVAL, "null Adam params");
  if (p->hna || p->ez_mode != XL_EZ_NONE || p->nin != 1 || p->ncomp != 1 || p->mask[0].kind != XL_MASK_SLM ||
      p->mask[1].kind != XL_MASK_NONE || p->gparam[0][1] || p->gparam[1][0] || p->gparam[1][1])
    return fail(XL_EINVAL, "the fused SLM adjoint + Adam step takes one component and one XL_MASK_SLM stage");
  if (ad->param != p->mask[0].m0) return fail(XL_EINVAL, "Adam param must be the SLM phase array (mask[0].m0)");
  if (p->mask[0].forms) return fail(XL_EINVAL, "the fused Adam step takes float32 phases (mask forms 0)");
  if (!ad->m || !ad->v || !ad->count) return fail(XL_EINVAL, "Adam m, v and count are required");
  if (p->batch > 1 && p->mask_batch_stride < (int64_t)p->rows * p->cols)
    return fail(XL_EINVAL, "a batched fused Adam step needs one phase array per member (mask_batch_stride >= "
                           "rows * cols): members sharing phases would race on them");
  if (!(ad->b1 >= 0.f && ad->b1 < 1.f && ad->b2 >= 0.f && ad->b2 < 1.f)) return fail(XL_EINVAL, "Adam betas must be in [0, 1)");
  xl::AdamArgs d;
  d.p = ad->param;
  d.g = nullptr;
  d.m = ad->m;
  d.v = ad->v;
  d.n = (int64_t)p->rows * p->cols;
  d.count = ad->count;
  d.lr = ad->lr;
  d.b1 = ad->b1;
  d.b2 = ad->b2;
  d.eps = ad->eps;
  d.wd = ad->weight_decay;
  d.vec = 0;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_PRE_BWD, st);
  const int64_t npix = (int64_t)p->rows * p->cols;
  hipLaunchKernelGGL(xl::pre_slm_backward_adam_kernel,
                     grid_for((npix + xl::kSlmUnroll - 1) / xl::kSlmUnroll, p->batch), dim3(256), 0, st, a, d);
  return hip_check(hipGetLastError(), "fused SLM adjoint + Adam launch");
}

int xl_adam_tick(int32_t* count, void* stream) {
  if (!count) return fail(XL_EINVAL, "null count");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(xl::adam_tick_kernel, dim3(1), dim3(64), 0, st, count);
  return hip_check(hipGetLastError(), "Adam tick launch");
}

size_t xl_mse_workspace_bytes(int64_t npix, int32_t batch) {
  if (npix < 1 || batch < 1) return 0;
  return (size_t)batch * xl::kMseBlocks * sizeof(float) + 256;
}

int xl_mse_intensity(const xl_mse_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->field || !p->target || p->npix < 1 || p->batch < 1 || p->batch > 65535)
    return fail(XL_EINVAL, "bad MSE arguments");
  if (!p->grad && !p->loss) return fail(XL_EINVAL, "MSE call asks for neither loss nor gradient");
  const size_t need = xl_mse_workspace_bytes(p->npix, p->batch);
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "MSE workspace too small: need " + std::to_string(need));
  xl::MseArgs a;
  a.field = (const float2*)p->field;
  a.fbs = p->field_batch_stride;
  a.target = p->target;
  a.grad = (float2*)p->grad;
  a.gbs = p->grad_batch_stride;
  a.partial = (float*)align256((char*)ws);
  a.loss = p->loss;
  a.npix = p->npix;
  a.gscale = (float)(4.0 * p->scale / (double)p->npix);
  a.lscale = (float)(p->scale / (double)p->npix);
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_MSE, st);
  hipLaunchKernelGGL(xl::mse_kernel, dim3(xl::kMseBlocks, p->batch), dim3(xl::kMseThreads), 0, st, a);
  if (p->loss)
    hipLaunchKernelGGL(xl::mse_reduce_kernel, dim3(p->batch), dim3(xl::kMseThreads), 0, st, a, xl::kMseBlocks);
  return hip_check(hipGetLastError(), "MSE launch");
}

int xl_adam_step(const xl_adam_params* p, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->param || !p->grad || !p->m || !p->v || !p->count || p->n < 1) return fail(XL_EINVAL, "bad Adam arguments");
  if (!(p->b1 >= 0.f && p->b1 < 1.f && p->b2 >= 0.f && p->b2 < 1.f)) return fail(XL_EINVAL, "Adam betas must be in [0, 1)");
  xl::AdamArgs a;
  a.p = p->param;
  a.g = p->grad;
  a.m = p->m;
  a.v = p->v;
  a.n = p->n;
  a.count = p->count;
  a.lr = p->lr;
  a.b1 = p->b1;
  a.b2 = p->b2;
  a.eps = p->eps;
  a.wd = p->weight_decay;
  auto al16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
  a.vec = al16(a.p) && al16(a.g) && al16(a.m) && al16(a.v);
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_ADAM, st);
  int64_t blocks = (p->n / 4 + 255) / 256;
  blocks = std::min<int64_t>(std::max<int64_t>(blocks, 1), 8192);
  hipLaunchKernelGGL(xl::adam_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
  hipLaunchKernelGGL(xl::adam_tick_kernel, dim3(1), dim3(64), 0, st, p->count);
  return hip_check(hipGetLastError(), "Adam launch");
}

// ---------------------------------------------------------------- losses
size_t xl_field_loss_workspace_bytes(int64_t npix, int32_t batch) {
  if (npix < 1 || batch < 1) return 0;
  return (size_t)batch * xl::kMseBlocks * sizeof(float) + 256;
}

int xl_field_loss(const xl_field_loss_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->field || !p->target || p->npix < 1 || p->batch < 1 || p->batch > 65535)
    return fail(XL_EINVAL, "bad field-loss arguments");
  if (p->mode < XL_LOSS_INTENSITY || p->mode > XL_LOSS_PHASE) return fail(XL_EINVAL, "unknown loss mode");
  if (!p->grad && !p->loss) return fail(XL_EINVAL, "loss call asks for neither loss nor gradient");
  const size_t need = xl_field_loss_workspace_bytes(p->npix, p->batch);
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "loss workspace too small: need " + std::to_string(need));
  xl::FieldLossArgs a;
  a.field = (const float2*)p->field;
  a.fbs = p->field_batch_stride;
  a.target = (const float2*)p->target;
  a.tbs = p->target_batch_stride;
  a.grad = (float2*)p->grad;
  a.gbs = p->grad_batch_stride;
  a.partial = (float*)align256((char*)ws);
  a.loss = p->loss;
  a.npix = p->npix;
  const double s = p->scale / (double)p->npix;
  a.gscale = (float)((p->mode == XL_LOSS_INTENSITY ? 4.0 : 2.0) * s);
  a.lscale = (float)s;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_LOSS, st);
  const dim3 g(xl::kMseBlocks, p->batch), b(xl::kMseThreads);
  if (p->mode == XL_LOSS_INTENSITY) hipLaunchKernelGGL(xl::field_loss_kernel<xl::LOSS_INTENSITY>, g, b, 0, st, a);
  else if (p->mode == XL_LOSS_AMPLITUDE) hipLaunchKernelGGL(xl::field_loss_kernel<xl::LOSS_AMPLITUDE>, g, b, 0, st, a);
  else hipLaunchKernelGGL(xl::field_loss_kernel<xl::LOSS_PHASE>, g, b, 0, st, a);
  if (p->loss)
    hipLaunchKernelGGL(xl::field_loss_reduce_kernel, dim3(p->batch), b, 0, st, a, xl::kMseBlocks);
  return hip_check(hipGetLastError(), "loss launch");
}

size_t xl_small_area_workspace_bytes(int64_t npix, int32_t batch) {
  if (npix < 1 || batch < 1) return 0;
  return (size_t)batch * (xl::kMseBlocks * 2 + 4) * sizeof(float) + 512;
}

int xl_small_area(const xl_small_area_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->intensity || !p->loss || p->npix < 1 || p->batch < 1 || p->batch > 65535)
    return fail(XL_EINVAL, "bad small-area arguments");
  const size_t need = xl_small_area_workspace_bytes(p->npix, p->batch);
  if (!ws || ws_bytes < need)
    return fail(XL_EWORKSPACE, "small-area workspace too small: need " + std::to_string(need));
  xl::SmallAreaArgs a;
  a.in = p->intensity;
  a.ibs = p->batch_stride;
  a.grad = p->grad;
  a.gbs = p->grad_batch_stride;
  a.partial = (float*)align256((char*)ws);
  a.stats = (float*)align256((char*)(a.partial + (size_t)p->batch * xl::kMseBlocks * 2));
  a.loss = p->loss;
  a.npix = p->npix;
  a.scale = p->scale;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_LOSS, st);
  const dim3 g(xl::kMseBlocks, p->batch), b(xl::kMseThreads);
  hipLaunchKernelGGL(xl::small_area_kernel<0>, g, b, 0, st, a);
  hipLaunchKernelGGL(xl::small_area_reduce_kernel<0>, dim3(p->batch), b, 0, st, a, xl::kMseBlocks);
  hipLaunchKernelGGL(xl::small_area_kernel<1>, g, b, 0, st, a);
  hipLaunchKernelGGL(xl::small_area_reduce_kernel<1>, dim3(p->batch), b, 0, st, a, xl::kMseBlocks);
  if (p->grad) hipLaunchKernelGGL(xl::small_area_grad_kernel, g, b, 0, st, a);
  return hip_check(hipGetLastError(), "small-area launch");
}

// ---------------------------------------------------------------- focal second moment
size_t xl_moment_workspace_bytes(int64_t npix) {
  if (npix < 1) return 0;
  return (size_t)xl::kMseBlocks * 2 * sizeof(float) + 256;
}

static int moment_args(const xl_moment_params* p, xl::MomentArgs* a) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->field || !p->r2 || !p->stats || p->npix < 1 || p->ncomp < 1 || p->ncomp > 3)
    return fail(XL_EINVAL, "bad focal-moment arguments");
  a->e = (const float2*)p->field;
  a->cs = p->comp_stride;
  a->nc = p->ncomp;
  a->r2 = p->r2;
  a->npix = p->npix;
  a->partial = nullptr;
  a->stats = p->stats;
  a->loss = p->loss;
  a->gscale = p->gscale;
  a->grad = (float2*)p->grad;
  return XL_OK;
}

int xl_moment_forward(const xl_moment_params* p, void* ws, size_t ws_bytes, void* stream) {
  xl::MomentArgs a;
  if (int rc = moment_args(p, &a)) return rc;
  const size_t need = xl_moment_workspace_bytes(p->npix);
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "moment workspace too small: need " + std::to_string(need));
  a.partial = (float*)align256((char*)ws);
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_LOSS, st);
  hipLaunchKernelGGL(xl::moment_kernel, dim3(xl::kMseBlocks), dim3(xl::kMseThreads), 0, st, a);
  hipLaunchKernelGGL(xl::moment_reduce_kernel, dim3(1), dim3(xl::kMseThreads), 0, st, a, xl::kMseBlocks);
  return hip_check(hipGetLastError(), "moment launch");
}

int xl_moment_backward(const xl_moment_params* p, void* stream) {
  xl::MomentArgs a;
  if (int rc = moment_args(p, &a)) return rc;
  if (!p->grad) return fail(XL_EINVAL, "null gradient output");
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_LOSS, st);
  hipLaunchKernelGGL(xl::moment_grad_kernel, grid_for(p->npix), dim3(256), 0, st, a);
  return hip_check(hipGetLastError(), "moment gradient launch");
}

int xl_quality_factor(const double* z_dev, double rmax, double dr_real, double wavelength, double* out,
                      void* stream) {
  if (!z_dev || !out || !(dr_real > 0.0)) return fail(XL_EINVAL, "bad quality-factor arguments");
  hipLaunchKernelGGL(xl::quality_factor_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, z_dev, rmax, dr_real,
                     wavelength, out);
  return hip_check(hipGetLastError(), "quality factor launch");
}

// ---------------------------------------------------------------- distance gradient
size_t xl_dz_dot_workspace_bytes(void) { return (size_t)xl::kDzBlocks * sizeof(double) + 256; }

int xl_dz_dot(const xl_dz_dot_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->out || p->rows < 1 || p->cols < 1 || p->ncomp < 0 || p->ncomp > 3 || (p->ncomp && (!p->d || !p->g)))
    return fail(XL_EINVAL, "bad distance-gradient arguments");
  if (!p->ncomp && !p->pre) return fail(XL_EINVAL, "ncomp = 0 needs pre (the Ez_in term)");
  if (p->pre && (!p->gez || !p->x || !p->y)) return fail(XL_EINVAL, "the Ez_in term needs gez, x and y");
  const size_t need = xl_dz_dot_workspace_bytes();
  if (!ws || ws_bytes < need) return fail(XL_EWORKSPACE, "dz workspace too small: need " + std::to_string(need));
  xl::DzDotArgs a;
  a.d = (const float2*)p->d;
  a.g = (const float2*)p->g;
  a.cs = p->comp_stride;
  a.nc = p->ncomp;
  a.pre = (const float2*)p->pre;
  a.pre_cs = p->pre_comp_stride;
  a.gez = (const float2*)p->gez;
  a.x = p->x;
  a.y = p->y;
  a.rows = p->rows;
  a.cols = p->cols;
  a.z = p->z;
  a.z_dev = p->z_dev;
  a.partial = (double*)align256((char*)ws);
  a.out = p->out;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_LOSS, st);
  hipLaunchKernelGGL(xl::dz_dot_kernel, dim3(xl::kDzBlocks), dim3(xl::kDzThreads), 0, st, a);
  hipLaunchKernelGGL(xl::dz_dot_reduce_kernel, dim3(1), dim3(xl::kDzThreads), 0, st, a, xl::kDzBlocks);
  return hip_check(hipGetLastError(), "dz dot launch");
}

// ---------------------------------------------------------------- beam splitters
int xl_beam_splitter(const xl_bs_params* p, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->a || p->npix < 1) return fail(XL_EINVAL, "beam splitter needs input a and npix >= 1");
  if (!p->c && !p->d) return fail(XL_EINVAL, "beam splitter call has no output");
  xl::BsArgs a;
  a.a = (const float2*)p->a;
  a.b = (const float2*)p->b;
  a.acs = p->a_comp_stride;
  a.bcs = p->b_comp_stride;
  a.c = (float2*)p->c;
  a.d = (float2*)p->d;
  a.ccs = p->c_comp_stride;
  a.dcs = p->d_comp_stride;
  a.ca = make_float2(p->ca[0], p->ca[1]);
  a.cb = make_float2(p->cb[0], p->cb[1]);
  a.da = make_float2(p->da[0], p->da[1]);
  a.db = make_float2(p->db[0], p->db[1]);
  a.npix = p->npix;
  a.zero_ez = p->zero_ez ? 1 : 0;
  a.coef = (const float2*)p->coef;
  if (p->theta_mode < 0 || p->theta_mode > 7 || p->theta_mode == XL_BS_THETA_RAW)
    return fail(XL_EINVAL, "theta_mode must be 0..3, 1..3 + XL_BS_THETA_RAW");
  if (p->theta_mode && !p->theta) return fail(XL_EINVAL, "theta_mode needs the device angle theta");
  a.theta = p->theta;
  a.theta_mode = p->theta_mode;
  a.adjoint = p->adjoint ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_BS, st);
  const int64_t blocks = std::min<int64_t>((p->npix + 255) / 256, 4096);
  hipLaunchKernelGGL(xl::bs_kernel, dim3((unsigned)blocks, a.zero_ez ? 3 : 2), dim3(256), 0, st, a);
  return hip_check(hipGetLastError(), "beam splitter launch");
}

size_t xl_bs_dots_workspace_bytes(int64_t npix) {
  if (npix < 1) return 0;
  return (size_t)xl::kBsDotBlocks * 8 * sizeof(float) + 256;
}

int xl_bs_dots(const xl_bs_dots_params* p, void* ws, size_t ws_bytes, void* stream) {
  if (!p) return fail(XL_EINVAL, "null params");
  if (!p->a || p->npix < 1 || (!p->dots && !p->dtheta)) return fail(XL_EINVAL, "bad beam-splitter dot arguments");
  if (p->theta_mode < 0 || p->theta_mode > 7 || p->theta_mode == XL_BS_THETA_RAW ||
      (p->theta_mode && (!p->theta || !p->dtheta)))
    return fail(XL_EINVAL, "theta_mode 1..3 (+ XL_BS_THETA_RAW) needs theta and dtheta");
  const size_t need = xl_bs_dots_workspace_bytes(p->npix);
  if (!ws || ws_bytes < need)
    return fail(XL_EWORKSPACE, "beam-splitter dot workspace too small: need " + std::to_string(need));
  xl::BsDotsArgs a;
  a.gc = (const float2*)p->gc;
  a.gd = (const float2*)p->gd;
  a.gccs = p->gc_comp_stride;
  a.gdcs = p->gd_comp_stride;
  a.a = (const float2*)p->a;
  a.b = (const float2*)p->b;
  a.acs = p->a_comp_stride;
  a.bcs = p->b_comp_stride;
  a.partial = (float*)align256((char*)ws);
  a.dots = p->dots;
  a.npix = p->npix;
  a.theta = p->theta;
  a.theta_mode = p->theta_mode;
  a.dtheta = p->dtheta;
  if ((p->ga || p->gb) && (!p->theta_mode || p->g_planes < 2 || p->g_planes > 3))
    return fail(XL_EINVAL, "the fused input cotangents need theta_mode and g_planes 2 or 3");
  a.ga = (float2*)p->ga;
  a.gb = (float2*)p->gb;
  a.gacs = p->ga_comp_stride;
  a.gbcs = p->gb_comp_stride;
  a.g_planes = p->g_planes;
  hipStream_t st = (hipStream_t)stream;
  xl::KTimer t(xl::K_BS, st);
  hipLaunchKernelGGL(xl::bs_dots_kernel, dim3(xl::kBsDotBlocks), dim3(256), 0, st, a);
  hipLaunchKernelGGL(xl::bs_dots_reduce_kernel, dim3(1), dim3(256), 0, st, a, xl::kBsDotBlocks);
  return hip_check(hipGetLastError(), "beam-splitter dot launch");
}

int xl_uniform_phases(float* out, const uint64_t* seeds, const int32_t* enable, int64_t count, int32_t batch,
                      int64_t batch_stride, void* stream) {
  if (!out || !seeds) return fail(XL_EINVAL, "null pointer argument");
  if (count < 1 || batch < 1 || batch > 65535) return fail(XL_EINVAL, "bad phase-mask count / batch");
  if (batch > 1 && batch_stride < count)
    return fail(XL_EINVAL, "batch_stride must be >= count (members write disjoint masks)");
  xl::PhaseArgs a;
  a.out = out;
  a.seeds = seeds;
  a.enable = enable;
  a.count = count;
  a.bs = batch > 1 ? batch_stride : 0;
  hipLaunchKernelGGL(xl::uniform_phase_kernel, grid_for(count, batch), dim3(256), 0, (hipStream_t)stream, a);
  return hip_check(hipGetLastError(), "phase launch");
}

int xl_transfer_function_rs(void* out, const double* x, const double* y, int32_t nx, int32_t ny, double z, double k,
                            void* stream) {
  if (!out || !x || !y || nx < 1 || ny < 1) return fail(XL_EINVAL, "bad transfer-function arguments");
  xl::KTimer t(xl::K_TF, (hipStream_t)stream);
  hipLaunchKernelGGL(xl::transfer_function_kernel, grid_for((int64_t)nx * ny), dim3(256), 0, (hipStream_t)stream,
                     (float2*)out, x, y, nx, ny, z, k);
  return hip_check(hipGetLastError(), "transfer function launch");
}

int xl_timing_enable(int32_t enable, int32_t capacity) {
  xl::TimingState& s = xl::timing_state();
  std::lock_guard<std::mutex> lk(s.mu);
  for (hipEvent_t e : s.pool) (void)hipEventDestroy(e);
  s.pool.clear();
  s.ids.clear();
  s.used = 0;
  s.dropped = 0;
  s.on = false;
  ++s.gen;
  if (!enable) return XL_OK;
  if (capacity < 1 || capacity > (1 << 20)) return fail(XL_EINVAL, "timing capacity must be 1..2^20");
  s.pool.resize(2 * (size_t)capacity);
  s.ids.resize(capacity);
  for (auto& e : s.pool)
    if (int rc = hip_check(hipEventCreate(&e), "hipEventCreate")) return rc;
  s.on = true;
  return XL_OK;
}

int xl_timing_collect(xl_kernel_time* out, int32_t max_entries) {
  xl::TimingState& s = xl::timing_state();
  std::lock_guard<std::mutex> lk(s.mu);
  double tot[xl::K_COUNT] = {};
  int64_t cnt[xl::K_COUNT] = {};
  for (size_t i = 0; i < s.used; ++i) {
    if (int rc = hip_check(hipEventSynchronize(s.pool[2 * i + 1]), "hipEventSynchronize")) return -rc;
    float ms = 0.f;
    if (int rc = hip_check(hipEventElapsedTime(&ms, s.pool[2 * i], s.pool[2 * i + 1]), "hipEventElapsedTime"))
      return -rc;
    tot[s.ids[i]] += ms;
    cnt[s.ids[i]] += 1;
  }
  s.used = 0;
  ++s.gen;
  int n = 0;
  for (int k = 0; k < xl::K_COUNT; ++k) {
    if (!cnt[k]) continue;
    if (out && n < max_entries) {
      std::memset(out[n].name, 0, sizeof(out[n].name));
      std::strncpy(out[n].name, xl::kernel_name(k), sizeof(out[n].name) - 1);
      out[n].launches = cnt[k];
      out[n].total_ms = tot[k];
    }
    ++n;
  }
  return n;
}

}  // extern "C"
