// Host runtime of libmesolf: error state, contexts, PSF preparation and
// residency, patches (reconstruction problems), the RL driver and the two
// drop-in entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"
#include "rl_mfma.h"
#include "state.h"

namespace mlf {

// ================================================================ errors
static thread_local std::string g_last_error;

mlf_status set_error(mlf_status s, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

mlf_status clear_error() {
  g_last_error.clear();
  return MLF_OK;
}

const char* last_error() { return g_last_error.c_str(); }

// ======================================================== MATLAB helpers
static long matlab_round(double x) {
  return (long)(std::floor(std::fabs(x) + 0.5)) * (x >= 0 ? 1 : -1);
}

// recon:156-188.  mode(Nnum)==0 is false for every N>0 (mode of a scalar is
// the scalar), so loop_n = (N+1)/2 and the loop runs to floor(loop_n).
std::vector<std::pair<int, int>> spiral_order(int N) {
  std::vector<std::pair<int, int>> out;
  out.reserve((size_t)N * N);
  const int loops = (N + 1) / 2;
  for (int k = 1; k <= loops; ++k) {
    const int cs = N - 2 * (k - 1);
    const int b = k - 1;
    if (cs > 1) {
      for (int i = 1; i <= cs; ++i) out.emplace_back(i - 1 + b, b);
      for (int j = 2; j <= cs; ++j) out.emplace_back(cs - 1 + b, j - 1 + b);
      for (int i = cs - 1; i >= 1; --i) out.emplace_back(i - 1 + b, cs - 1 + b);
      for (int j = cs - 1; j >= 2; --j) out.emplace_back(b, j - 1 + b);
    } else {
      out.emplace_back(b, b);
    }
  }
  return out;
}

// recon:30-38 on per-view sums (NaN -> 0, /sum, -min, /max, <cut -> 0, *0.3).
std::vector<double> view_weights(const std::vector<double>& sums, double cut) {
  std::vector<double> w(sums);
  for (double& x : w)
    if (std::isnan(x)) x = 0.0;
  double s = 0.0;
  for (double x : w) s += x;
  for (double& x : w) x /= s;
  double mn = w[0];
  for (double x : w) mn = std::min(mn, x);
  for (double& x : w) x -= mn;
  double mx = w[0];
  for (double x : w) mx = std::max(mx, x);
  for (double& x : w) x /= mx;
  for (double& x : w)
    if (x < cut) x = 0.0;
  for (double& x : w) x *= 0.3;
  return w;
}

// recon:20-26: NIP_range = ceil(Z/3) rounded up to even; 3*NIP_range filters.
int nip_range(int Z) {
  int r = (Z + 2) / 3;
  return (r % 2 == 1) ? r + 1 : r;
}

// sample_filter plane -> 0-based filter index (or -1), recon:224-231.
std::vector<int> filter_plane_index(int Z) {
  const int nf = 3 * nip_range(Z);
  const long half_z = matlab_round(Z / 2.0);
  std::vector<int> idx(Z, -1);
  for (int z = 1; z <= Z; ++z) {
    const double lo = half_z - nf / 2.0 + 1.0, hi = half_z + nf / 2.0;
    if (z >= lo && z <= hi) idx[z - 1] = (int)(z - half_z + nf / 2) - 1;
  }
  return idx;
}

double filter_omega(int Z, int N, int f) {
  const int nf = 3 * nip_range(Z);
  const double axis = (f + 1) - nf / 2.0;
  double om = (N - 7) * std::exp(-std::pow(axis / 100.0 * 2.0 * std::sqrt(std::log(2.0)), 2.0));
  return om < 0.2 ? 0.2 : om;
}

// Real-space circular kernel of one axis of the filter bank (recon:54-60):
// kappa = real(ifft(|fft(g)| / max|fft(g)|)), g(x) = exp(-((x - I/2)/omega)^2),
// x = 1..I.  Direct DFT in FP64 with exact (k*n mod I) twiddle indexing.
std::vector<double> filter_kernel_1d(int I, double omega) {
  std::vector<double> g(I), c(I), s(I), a(I), k(I);
  const double two_pi = 6.283185307179586476925286766559;
  for (int m = 0; m < I; ++m) {
    c[m] = std::cos(two_pi * m / I);
    s[m] = std::sin(two_pi * m / I);
  }
  for (int n = 0; n < I; ++n) {
    const double x = (n + 1) - I / 2.0;
    g[n] = std::exp(-(x / omega) * (x / omega));
  }
  double mx = 0.0;
  for (int f = 0; f < I; ++f) {
    double re = 0.0, im = 0.0;
    long m = 0;
    for (int n = 0; n < I; ++n) {
      re += g[n] * c[m];
      im -= g[n] * s[m];
      m += f;
      if (m >= I) m -= I;
    }
    a[f] = std::sqrt(re * re + im * im);
    mx = std::max(mx, a[f]);
  }
  for (int f = 0; f < I; ++f) a[f] /= mx;
  for (int n = 0; n < I; ++n) {
    double re = 0.0;
    long m = 0;
    for (int f = 0; f < I; ++f) {
      re += a[f] * c[m];
      m += n;
      if (m >= I) m -= I;
    }
    k[n] = re / I;
  }
  return k;
}

// Truncate kappa (circular) to the smallest symmetric support whose dropped
// mass is <= 1e-10 of the total; returns (n0, taps) with taps[j] = kappa[n0+j].
static void truncate_kernel(const std::vector<double>& k, int* n0, std::vector<float>* taps) {
  const int I = (int)k.size();
  double total = 0.0;
  for (double v : k) total += std::fabs(v);
  const int half = (I - 1) / 2;
  int T = half;
  double kept = std::fabs(k[0]);
  for (int t = 1; t <= half; ++t) {
    kept += std::fabs(k[t]) + std::fabs(k[I - t]);
    if (total - kept <= 1e-10 * total) {
      T = t;
      break;
    }
  }
  taps->clear();
  if (2 * T + 1 >= I) {  // full period
    *n0 = -((I - 1) / 2);
    for (int j = 0; j < I; ++j) taps->push_back((float)k[((*n0 + j) % I + I) % I]);
  } else {
    *n0 = -T;
    for (int j = 0; j < 2 * T + 1; ++j) taps->push_back((float)k[((*n0 + j) % I + I) % I]);
  }
}

// ================================================================ context
}  // namespace mlf

using namespace mlf;

// Handle lifetimes: a PSF keeps its context alive and a patch (or batch)
// keeps its PSF alive, so destroying a parent first only defers its release.
std::mutex mlf::g_life_mu;

void mlf::ctx_free(mlf_context ctx) {
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

mlf_status mlf::ctx_activate(mlf_context ctx) {
  MLF_HIP(hipSetDevice(ctx->device));
  return MLF_OK;
}

extern "C" {

const char* mlf_last_error(void) { return mlf::last_error(); }

const char* mlf_status_string(mlf_status s) {
  switch (s) {
    case MLF_OK: return "ok";
    case MLF_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MLF_ERR_UNSUPPORTED: return "unsupported";
    case MLF_ERR_DEVICE: return "device error";
    case MLF_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MLF_ERR_BUDGET: return "HBM budget exceeded";
    case MLF_ERR_INTERNAL: return "internal error";
    case MLF_ERR_CANCELLED: return "cancelled";
    case MLF_ERR_IO: return "file error";
  }
  return "unknown status";
}

int32_t mlf_abi_version(void) { return MLF_ABI_VERSION; }

mlf_status mlf_context_create(int32_t device, mlf_context* out) {
  clear_error();
  if (!out) return set_error(MLF_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return set_error(MLF_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "device %d out of range [0, %d)", device, count);
  auto ctx = std::make_unique<mlf_context_s>();
  ctx->device = device;
  MLF_HIP(hipSetDevice(device));
  MLF_CHECK(noalloc_scratch_init(device));
  MLF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  *out = ctx.release();
  return MLF_OK;
}

mlf_status mlf_context_destroy(mlf_context ctx) {
  clear_error();
  if (!ctx) return MLF_OK;
  std::lock_guard<std::mutex> lk(g_life_mu);
  if (ctx->closed) return set_error(MLF_ERR_INVALID_ARGUMENT, "context destroyed twice");
  ctx->closed = true;
  if (ctx->refs == 0) ctx_free(ctx);
  return MLF_OK;
}

mlf_status mlf_context_synchronize(mlf_context ctx) {
  clear_error();
  if (!ctx) return set_error(MLF_ERR_INVALID_ARGUMENT, "ctx is NULL");
  MLF_CHECK(ctx_activate(ctx));
  MLF_HIP(hipStreamSynchronize(ctx->stream));
  return MLF_OK;
}

mlf_status mlf_context_memory(mlf_context ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  clear_error();
  if (!ctx) return set_error(MLF_ERR_INVALID_ARGUMENT, "ctx is NULL");
  MLF_CHECK(ctx_activate(ctx));
  size_t f = 0, t = 0;
  MLF_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return MLF_OK;
}

}  // extern "C"

// =================================================================== PSF

// Called with g_life_mu held.
void mlf::psf_free(mlf_psf psf) {
  mlf_context ctx = psf->ctx;
  (void)hipSetDevice(ctx->device);
  delete psf;
  if (--ctx->refs == 0 && ctx->closed) ctx_free(ctx);
}

static int64_t plane_block(int r) { return (int64_t)(2 * r + 1) * (2 * r + 1) * kP; }


// Build the packed forward/backward tables of one precision (once per PSF).
static mlf_status psf_ensure_pack(mlf_psf_s* ps, Precision prec) {
  PrecPack& pk = ps->pack[(int)prec];
  if (pk.built) return MLF_OK;
  const int Z = ps->Z, V = ps->V;
  const int phi = fwd_phi(prec);
  pk.woff.assign(Z, 0);
  pk.boff.assign(Z, 0);
  pk.nchunk.assign(Z, 0);
  pk.epad.assign(Z, 0);
  std::vector<int32_t> rows_f(Z), rows_b(Z);
  int rfmax = 0;
  for (int r : ps->rf) rfmax = std::max(rfmax, r);
  pk.tpz = rfmax <= tpz_max_radius();
  int64_t of = 0, ob = 0;
  for (int z = 0; z < Z; ++z) {
    const int nt_f = (2 * ps->rf[z] + 1) * (2 * ps->rf[z] + 1);
    const int nt_b = (2 * ps->rb[z] + 1) * (2 * ps->rb[z] + 1);
    pk.nchunk[z] = (nt_f + kFwdChunk - 1) / kFwdChunk;
    const int Kf = 2 * ps->rf[z] + 1;
    rows_f[z] = pk.tpz ? Kf * 16 * ((Kf + 15) / 16) : pk.nchunk[z] * kFwdChunk;
    const int Kb = 2 * ps->rb[z] + 1;
    const int nt_b_pad = prec != Precision::kF32 ? 32 * bwd_quad_steps(Kb) : nt_b;
    const int tpad = prec != Precision::kF32 ? kBwdTapPad16 : kBwdTapPad;
    pk.epad[z] = (nt_b_pad + tpad - 1) / tpad * tpad;
    rows_b[z] = pk.epad[z];
    pk.woff[z] = of;
    of += (int64_t)rows_f[z] * phi;
    pk.boff[z] = ob;
    ob += (int64_t)kPhiPadBwd * pk.epad[z];
  }
  pk.pvf = of;
  pk.pvb = ob;
  const size_t es = elem_bytes(prec);
  MLF_CHECK(pk.fwd.alloc((size_t)of * V * es));
  MLF_CHECK(pk.bwd.alloc((size_t)ob * V * es));
  MLF_CHECK(upload_vec(pk.d_woff, pk.woff));
  MLF_CHECK(upload_vec(pk.d_boff, pk.boff));
  MLF_CHECK(upload_vec(pk.d_nchunk, pk.nchunk));
  MLF_CHECK(upload_vec(pk.d_epad, pk.epad));
  DevBuf<int32_t> d_rows_f, d_rows_b;
  MLF_CHECK(upload_vec(d_rows_f, rows_f));
  MLF_CHECK(upload_vec(d_rows_b, rows_b));
  hipStream_t st = ps->ctx->stream;
  if (prec == Precision::kF16) {
    // pre-scale every (view, plane) block so its largest tap is 1
    if (!pk.tpz) return set_error(MLF_ERR_UNSUPPORTED, "FP16 needs PSF radii <= %d", tpz_max_radius());
    MLF_CHECK(pk.d_fscale.alloc((size_t)V * Z));
    MLF_CHECK(pk.d_finv.alloc((size_t)V * Z));
    MLF_CHECK(pk.d_bscale.alloc((size_t)V * Z));
    MLF_LAUNCH(launch_block_scale(ps->fwd.ptr, ps->d_offf.ptr, ps->d_rf.ptr, ps->per_view_f, Z, V,
                                  pk.d_fscale.ptr, pk.d_finv.ptr, ps->ctx->stream));
    MLF_LAUNCH(launch_block_scale(ps->bwd.ptr, ps->d_offb.ptr, ps->d_rb.ptr, ps->per_view_b, Z, V,
                                  pk.d_bscale.ptr, nullptr, ps->ctx->stream));
  }
  PackGeo gf{ps->d_rf.ptr, d_rows_f.ptr, ps->d_offf.ptr, pk.d_woff.ptr, ps->per_view_f, pk.pvf, 0,
             pk.d_fscale.ptr};
  hipError_t e = pk.tpz ? launch_pack_fwd_tpz(prec, ps->fwd.ptr, pk.fwd.ptr, gf, Z, V, st)
                        : launch_pack_fwd(prec, ps->fwd.ptr, pk.fwd.ptr, gf, Z, V, st);
  if (e != hipSuccess) return set_error(MLF_ERR_DEVICE, "pack_fwd: %s", hipGetErrorString(e));
  PackGeo gb{ps->d_rb.ptr, d_rows_b.ptr, ps->d_offb.ptr, pk.d_boff.ptr, ps->per_view_b, pk.pvb,
             prec != Precision::kF32 ? 1 : 0, pk.d_bscale.ptr};
  e = launch_pack_bwd(prec, ps->bwd.ptr, pk.bwd.ptr, gb, Z, V, st);
  if (e != hipSuccess) return set_error(MLF_ERR_DEVICE, "pack_bwd: %s", hipGetErrorString(e));
  MLF_HIP(hipStreamSynchronize(st));
  pk.built = true;
  return MLF_OK;
}

// Normaliser table for a grid of n1 x n2 lenslets (shared by every grid that
// is at least 2*rb_max+1 lenslets along an axis).
static mlf_status psf_ensure_norm(mlf_psf_s* ps, Precision prec, int n1, int n2, NormTable** out) {
  int rbmax = 0;
  for (int r : ps->rb) rbmax = std::max(rbmax, r);
  const int key1 = n1 >= 2 * rbmax + 1 ? -1 : n1;
  const int key2 = n2 >= 2 * rbmax + 1 ? -1 : n2;
  for (auto& t : ps->norms[(int)prec])
    if (t->key1 == key1 && t->key2 == key2) {
      *out = t.get();
      return MLF_OK;
    }
  auto nt = std::make_unique<NormTable>();
  nt->key1 = key1;
  nt->key2 = key2;
  const int Z = ps->Z;
  // representative grid sizes: any size >= 2*rb_max+1 gives the same classes
  const int g1 = key1 < 0 ? 2 * rbmax + 1 : n1, g2 = key2 < 0 ? 2 * rbmax + 1 : n2;
  std::vector<int64_t> soff(Z), toff(Z);
  int64_t so = 0, to = 0;
  nt->C1.resize(Z);
  nt->C2.resize(Z);
  for (int z = 0; z < Z; ++z) {
    const int r = ps->rb[z];
    nt->C1[z] = norm_classes(g1, r);
    nt->C2[z] = norm_classes(g2, r);
    soff[z] = so;
    so += (int64_t)(2 * r + 1) * nt->C1[z] * kP;
    toff[z] = to;
    to += (int64_t)nt->C2[z] * nt->C1[z] * kNormRow;
  }
  nt->toff = toff;
  nt->pvt = to;
  MLF_CHECK(nt->T.alloc((size_t)to * ps->V));
  MLF_CHECK(upload_vec(nt->d_toff, toff));
  MLF_CHECK(upload_vec(nt->d_C1, nt->C1));
  DevBuf<int32_t> d_C2;
  DevBuf<int64_t> d_soff;
  MLF_CHECK(upload_vec(d_C2, nt->C2));
  MLF_CHECK(upload_vec(d_soff, soff));
  DevBuf<double> S;
  MLF_CHECK(S.alloc((size_t)so));
  NormGeo ng{g1, g2, ps->d_rb.ptr, nt->d_C1.ptr, d_C2.ptr, ps->d_offb.ptr, d_soff.ptr, nt->d_toff.ptr,
             nullptr};
  hipStream_t st = ps->ctx->stream;
  const float* bscale = ps->pack[(int)prec].d_bscale.ptr;  // FP16: taps as packed
  for (int v = 0; v < ps->V; ++v) {
    ng.scale = bscale ? bscale + (size_t)v * Z : nullptr;
    hipError_t e = launch_norm_table(prec, ps->bwd.ptr + (size_t)v * ps->per_view_b, S.ptr,
                                     nt->T.ptr + (size_t)v * to, ng, Z, st);
    if (e != hipSuccess) return set_error(MLF_ERR_DEVICE, "norm table: %s", hipGetErrorString(e));
  }
  MLF_HIP(hipStreamSynchronize(st));
  *out = nt.get();
  ps->norms[(int)prec].push_back(std::move(nt));
  return MLF_OK;
}

// Upload compact tables (host layout == device layout for the FP32 kernels).
static mlf_status upload_large(void* dst, const void* src, size_t bytes, hipStream_t st);

static mlf_status psf_upload(mlf_psf_s* p, const float* fwd, const float* bwd) {
  p->offf.assign(p->Z, 0);
  p->offb.assign(p->Z, 0);
  int64_t of = 0, ob = 0;
  for (int z = 0; z < p->Z; ++z) {
    p->offf[z] = of;
    of += plane_block(p->rf[z]);
    p->offb[z] = ob;
    ob += plane_block(p->rb[z]);
  }
  p->per_view_f = of;
  p->per_view_b = ob;
  MLF_CHECK(p->fwd.alloc((size_t)of * p->V));
  MLF_CHECK(p->bwd.alloc((size_t)ob * p->V));
  MLF_CHECK(upload_large(p->fwd.ptr, fwd, p->fwd.bytes(), p->ctx->stream));
  MLF_CHECK(upload_large(p->bwd.ptr, bwd, p->bwd.bytes(), p->ctx->stream));
  MLF_CHECK(p->d_rf.alloc(p->Z));
  MLF_CHECK(p->d_rb.alloc(p->Z));
  MLF_CHECK(p->d_offf.alloc(p->Z));
  MLF_CHECK(p->d_offb.alloc(p->Z));
  MLF_HIP(hipMemcpy(p->d_rf.ptr, p->rf.data(), p->d_rf.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->d_rb.ptr, p->rb.data(), p->d_rb.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->d_offf.ptr, p->offf.data(), p->d_offf.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->d_offb.ptr, p->offb.data(), p->d_offb.bytes(), hipMemcpyHostToDevice));
  return MLF_OK;
}

// Run f(i) for i in [0, n) on up to hardware_concurrency threads (static
// interleaved split, so every thread gets a mix of large and small items).
template <typename F>
static void parallel_for(int64_t n, F&& f) {
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(n, std::min(64u, std::max(1u, std::thread::hardware_concurrency()))));
  if (nt == 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int64_t i = t; i < n; i += nt) f(i);
    });
  for (auto& x : th) x.join();
}

// Host -> device copy of a large pageable buffer: 128 MB chunks are copied
// into two pinned staging buffers by all host threads and sent by DMA while
// the next chunk is staged (the runtime's own pageable path stages with one
// thread).  Synchronous: returns when the data is on the device.
static mlf_status upload_large(void* dst, const void* src, size_t bytes, hipStream_t st) {
  constexpr size_t kStage = (size_t)128 << 20;
  if (bytes < 2 * kStage) {
    MLF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MLF_OK;
  }
  struct Staging {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Staging() {
      for (int i = 0; i < 2; ++i) {
        if (ev[i]) (void)hipEventSynchronize(ev[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        if (buf[i]) (void)hipHostFree(buf[i]);
      }
    }
  } sg;
  for (int i = 0; i < 2; ++i) {
    MLF_HIP(hipHostMalloc(&sg.buf[i], kStage, hipHostMallocDefault));
    MLF_HIP(hipEventCreateWithFlags(&sg.ev[i], hipEventDisableTiming));
  }
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  for (size_t off = 0, i = 0; off < bytes; off += kStage, ++i) {
    const int b = (int)(i % 2);
    const size_t len = std::min(kStage, bytes - off);
    if (i >= 2) MLF_HIP(hipEventSynchronize(sg.ev[b]));  // its previous DMA finished
    constexpr size_t kPiece = (size_t)8 << 20;
    char* pb = static_cast<char*>(sg.buf[b]);
    parallel_for((int64_t)((len + kPiece - 1) / kPiece), [&](int64_t q) {
      const size_t o = (size_t)q * kPiece;
      std::memcpy(pb + o, s + off + o, std::min(kPiece, len - o));
    });
    MLF_HIP(hipMemcpyAsync(d + off, pb, len, hipMemcpyHostToDevice, st));
    MLF_HIP(hipEventRecord(sg.ev[b], st));
  }
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

// 64-bit content hash of n floats: per-chunk hashes (parallel) folded in
// chunk order.  Used only to recognise an identical PSF between calls.
static uint64_t content_hash(const float* p, size_t n) {
  constexpr size_t kChunk = (size_t)1 << 22;
  const size_t nc = (n + kChunk - 1) / kChunk;
  std::vector<uint64_t> hc(nc);
  parallel_for((int64_t)nc, [&](int64_t c) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p) + (size_t)c * kChunk;
    const size_t m = std::min(kChunk, n - (size_t)c * kChunk);
    uint64_t h0 = 0x9E3779B97F4A7C15ull ^ (uint64_t)c, h1 = 0xC2B2AE3D27D4EB4Full, h2 = 0x165667B19E3779F9ull,
             h3 = 0x27D4EB2F165667C5ull;
    size_t i = 0;
    for (; i + 4 <= m; i += 4) {
      h0 = (h0 ^ w[i]) * 0x100000001B3ull;
      h1 = (h1 ^ w[i + 1]) * 0x100000001B3ull;
      h2 = (h2 ^ w[i + 2]) * 0x100000001B3ull;
      h3 = (h3 ^ w[i + 3]) * 0x100000001B3ull;
    }
    for (; i < m; ++i) h0 = (h0 ^ w[i]) * 0x100000001B3ull;
    hc[c] = h0 ^ (h1 << 1 | h1 >> 63) ^ (h2 << 2 | h2 >> 62) ^ (h3 << 3 | h3 >> 61);
  });
  uint64_t h = 0xCBF29CE484222325ull ^ n;
  for (uint64_t x : hc) h = (h ^ x) * 0x100000001B3ull + (h >> 29);
  return h;
}

// Exact adjoint of a forward table: bwd[d] = fwd[-d] per plane.
static void adjoint_tables(const std::vector<int32_t>& rf, int V, const float* fwd,
                           std::vector<float>* bwd) {
  int64_t per_view = 0;
  for (int r : rf) per_view += plane_block(r);
  bwd->resize((size_t)per_view * V);
  for (int k = 0; k < V; ++k) {
    int64_t off = (int64_t)k * per_view;
    for (size_t z = 0; z < rf.size(); ++z) {
      const int r = rf[z], K = 2 * r + 1;
      for (int e2 = 0; e2 < K; ++e2)
        for (int e1 = 0; e1 < K; ++e1) {
          const float* src = fwd + off + ((int64_t)(K - 1 - e2) * K + (K - 1 - e1)) * kP;
          float* dst = bwd->data() + off + ((int64_t)e2 * K + e1) * kP;
          std::memcpy(dst, src, sizeof(float) * kP);
        }
      off += plane_block(r);
    }
  }
}

extern "C" {

mlf_status mlf_psf_create_compact(mlf_context ctx, int64_t N, int64_t Z, int64_t nviews,
                                  const int32_t* view_uv, const double* view_weight,
                                  const int32_t* rf, const float* fwd, const int32_t* rb,
                                  const float* bwd, mlf_psf* out) {
  clear_error();
  if (!ctx || !out || !view_uv || !view_weight || !rf || !fwd)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (N != kN) return set_error(MLF_ERR_UNSUPPORTED, "Nnum must be 15 (got %lld)", (long long)N);
  if (Z < 1 || Z > 4096) return set_error(MLF_ERR_INVALID_ARGUMENT, "bad plane count %lld", (long long)Z);
  if (nviews < 1 || nviews > kP)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad view count %lld", (long long)nviews);
  for (int64_t z = 0; z < Z; ++z) {
    if (rf[z] < 0 || rf[z] > 512 || (rb && (rb[z] < 0 || rb[z] > 512)))
      return set_error(MLF_ERR_INVALID_ARGUMENT, "bad radius at plane %lld", (long long)z);
  }
  for (int64_t k = 0; k < nviews; ++k) {
    if (view_uv[2 * k] < 0 || view_uv[2 * k] >= kN || view_uv[2 * k + 1] < 0 ||
        view_uv[2 * k + 1] >= kN)
      return set_error(MLF_ERR_INVALID_ARGUMENT, "view %lld has (u,v) outside [0,15)", (long long)k);
    if (!std::isfinite(view_weight[k]) || view_weight[k] == 0.0)
      return set_error(MLF_ERR_INVALID_ARGUMENT, "view %lld has weight %g", (long long)k,
                       view_weight[k]);
  }
  MLF_CHECK(ctx_activate(ctx));
  auto p = std::make_unique<mlf_psf_s>();
  p->ctx = ctx;
  {
    std::lock_guard<std::mutex> lk(g_life_mu);
    if (ctx->closed) return set_error(MLF_ERR_INVALID_ARGUMENT, "context already destroyed");
    ++ctx->refs;
  }
  struct RefUndo {  // drop the reference if creation fails below
    mlf_psf_s* p;
    ~RefUndo() {
      if (p) {
        std::lock_guard<std::mutex> lk(g_life_mu);
        --p->ctx->refs;
      }
    }
  } undo{p.get()};
  p->Z = (int)Z;
  p->V = (int)nviews;
  p->uv.assign(view_uv, view_uv + 2 * nviews);
  p->weight.assign(view_weight, view_weight + nviews);
  p->rf.assign(rf, rf + Z);
  std::vector<float> adj;
  if (bwd) {
    if (!rb) return set_error(MLF_ERR_INVALID_ARGUMENT, "bwd given without rb");
    p->rb.assign(rb, rb + Z);
  } else {
    p->rb = p->rf;
    adjoint_tables(p->rf, p->V, fwd, &adj);
    bwd = adj.data();
  }
  MLF_CHECK(psf_upload(p.get(), fwd, bwd));
  undo.p = nullptr;
  *out = p.release();
  return MLF_OK;
}

// Dense raw PSFs -> compact tables, following recon:13-40 and
// gen_phasespace_PSF_theory.m.  With Hcrop the cropped H (index i0 + x) and
// c0 = 7, the phase-space forward table of view (u,v) is
//   fwd[d1,d2,s1,s2] = Hn[i0 + u + 15(d1+R) - s1 + c0, i0 + v + 15(d2+R) - s2 + c0, s1, s2, z]
// (closed form of the de-shift + reshape/permute chain), and the backward
// table reads the cropped Ht at the view's own (u, v) slot (recon:121):
//   bwd[d1,d2,s1,s2] = Htn[i0 + 15(d1+R) + c0 + s1 - u, i0 + 15(d2+R) + c0 + s2 - v, u, v, z].
// Every raw element Hcrop[x, y, a, b, z] therefore belongs to exactly one
// view, u = (x + a - c0) mod 15, at lenslet offset d1 = (x + a - c0) / 15 - R
// (same along y), so the view weights (recon:30: W(u,v) is the sum of
// H_ps(:,:,u,v,:)) and each view's extent per plane come from one streaming
// pass over H; likewise for Ht, whose (u, v) slot is the view.  The passes run
// in parallel over planes (each plane's sums in a fixed order, planes added
// in order); the forward tables fill in parallel over (plane, v, d2), each
// task writing whole table rows, the backward tables over (view, plane).
mlf_status mlf_psf_create_dense(mlf_context ctx, const float* H, const float* Ht, int64_t S,
                                int64_t N, int64_t Z, mlf_psf* out) {
  clear_error();
  if (!ctx || !H || !Ht || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (N != kN) return set_error(MLF_ERR_UNSUPPORTED, "Nnum must be 15 (got %lld)", (long long)N);
  if (Z < 1) return set_error(MLF_ERR_INVALID_ARGUMENT, "Z must be >= 1");
  if (S < N) return set_error(MLF_ERR_INVALID_ARGUMENT, "PSF size %lld smaller than Nnum", (long long)S);
  const int64_t IMG = S - ((S - N) % (2 * N));  // gen_phasespace:6
  const int64_t i0 = matlab_round(S / 2.0) - (int64_t)std::trunc(IMG / 2.0) - 1;
  const int R = (int)((IMG / N - 1) / 2);
  const size_t SS = (size_t)S * S;
  const size_t slab = SS * N * N;
  auto plane_of = [&](const float* A, int a, int b, int64_t z) {
    return A + SS * ((size_t)a + (size_t)N * ((size_t)b + (size_t)N * (size_t)z));
  };
  // One streaming pass per (array, plane): the plane's sum over the whole
  // S x S x 15 x 15 slab (recon:13-16, FP64, in a fixed order) and, over the
  // cropped region, per view the raw sum of H_ps(:,:,u,v,z) and the largest
  // lenslet offset holding a non-zero tap (forward from H, backward from Ht).
  // Each row is read once; the cropped work reuses it from cache.
  std::vector<double> hs(Z, 0.0), hts(Z, 0.0);
  std::vector<double> wz((size_t)Z * kP, 0.0);  // raw (unnormalised) view sums of H per plane
  std::vector<int32_t> rfz((size_t)Z * kP, -1), rbz((size_t)Z * kP, -1);
  auto floor15 = [](int64_t t) { return t >= 0 ? t / kN : -((-t + kN - 1) / kN); };
  parallel_for(2 * Z, [&](int64_t q) {
    const bool isH = q < Z;
    const int64_t z = q % Z;
    const float* A = isH ? H : Ht;
    double* w = wz.data() + (size_t)z * kP;
    int32_t* ext = (isH ? rfz : rbz).data() + (size_t)z * kP;
    double total = 0.0;
    for (int b = 0; b < kN; ++b)
      for (int a = 0; a < kN; ++a) {
        const float* P = plane_of(A, a, b, z);
        for (int64_t y = 0; y < S; ++y) {
          const float* row = P + (size_t)y * S;
          double s4[4] = {0.0, 0.0, 0.0, 0.0};
          int64_t x = 0;
          for (; x + 4 <= S; x += 4)
            for (int j = 0; j < 4; ++j) s4[j] += row[x + j];
          for (; x < S; ++x) s4[0] += row[x];
          total += (s4[0] + s4[1]) + (s4[2] + s4[3]);
          const int64_t yc = y - i0;  // row of the cropped PSF
          if (yc < 0 || yc >= IMG) continue;
          const float* rc = row + i0;
          if (isH) {
            // H_ps(:,:,u,v,z) takes Hcrop(x, y, a, b) with xp = x + a - c0 =
            // 15 (d1 + R) + u, yp = y + b - c0 = 15 (d2 + R) + v
            const int64_t yp = yc + b - kC0;
            if (yp < 0 || yp >= IMG) continue;
            const int v = (int)(yp % kN), ad2 = std::abs((int)(yp / kN) - R);
            double acc[kN] = {0.0};
            int32_t m[kN];
            for (int u = 0; u < kN; ++u) m[u] = -1;
            const int64_t xp_lo = std::max<int64_t>(0, a - kC0), xp_hi = std::min<int64_t>(IMG, IMG + a - kC0);
            const float* src = rc + (kC0 - a);  // src[xp] = Hcrop(xp - a + c0, y)
            for (int64_t L = xp_lo / kN; L * kN < xp_hi; ++L) {
              const int dl = std::abs((int)L - R);
              const int64_t x0 = kN * L;
              if (x0 >= xp_lo && x0 + kN <= xp_hi) {  // a whole lenslet: fixed trip count
                unsigned nz = 0;
                for (int u = 0; u < kN; ++u) {
                  const float val = src[x0 + u];
                  acc[u] += val;
                  nz |= (unsigned)(val != 0.f) << u;
                }
                if (nz)
                  for (int u = 0; u < kN; ++u)
                    if (nz >> u & 1u) m[u] = std::max(m[u], dl);
                continue;
              }
              for (int64_t xp = std::max(x0, xp_lo); xp < std::min(x0 + kN, xp_hi); ++xp) {
                const int u = (int)(xp - x0);
                const float val = src[xp];
                acc[u] += val;
                if (val != 0.f) m[u] = std::max(m[u], dl);
              }
            }
            for (int u = 0; u < kN; ++u) {
              w[u + kN * v] += acc[u];
              if (m[u] >= 0) ext[u + kN * v] = std::max(ext[u + kN * v], std::max(m[u], ad2));
            }
          } else {
            // Ht's (u, v) slot is the view: d1 + R = floor((x - c0 + u) / 15),
            // monotone in x, so the row's first and last non-zero bound |d1|
            const int u = a, v = b;
            int64_t x0 = 0, x1 = IMG - 1;
            while (x0 < IMG && rc[x0] == 0.f) ++x0;
            if (x0 == IMG) continue;
            while (rc[x1] == 0.f) --x1;
            const int d2 = (int)floor15(yc - kC0 + v) - R;
            const int d1a = (int)floor15(x0 - kC0 + u) - R, d1b = (int)floor15(x1 - kC0 + u) - R;
            const int mm = std::max(std::max(std::abs(d1a), std::abs(d1b)), std::abs(d2));
            ext[u + kN * v] = std::max(ext[u + kN * v], mm);
          }
        }
      }
    (isH ? hs : hts)[z] = total;
  });
  // weights (recon:30): W(u,v) = sum over z of the view's sum in H_ps / hs(z)
  for (int64_t z = 0; z < Z; ++z)
    for (int k = 0; k < kP; ++k) wz[(size_t)z * kP + k] /= hs[z];
  std::vector<double> wsum(kP, 0.0);
  for (int64_t z = 0; z < Z; ++z)
    for (int k = 0; k < kP; ++k) wsum[k] += wz[(size_t)z * kP + k];
  // weights indexed [u + 15 v]
  std::vector<double> w = view_weights(wsum, 0.2);
  std::vector<int32_t> uv;
  std::vector<double> wv;
  for (auto& [u, v] : spiral_order(kN)) {
    const double x = w[u + kN * v];
    if (std::isnan(x))
      return set_error(MLF_ERR_INVALID_ARGUMENT,
                       "degenerate view weights (NaN): all views have equal PSF sums");
    if (x == 0.0) continue;
    uv.push_back(u);
    uv.push_back(v);
    wv.push_back(x);
  }
  const int V = (int)wv.size();
  if (V == 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "no active view");
  // per-plane radii: the widest active view
  std::vector<int32_t> rf(Z, 0), rb(Z, 0);
  for (int64_t z = 0; z < Z; ++z)
    for (int k = 0; k < V; ++k) {
      const int q = uv[2 * k] + kN * uv[2 * k + 1];
      rf[z] = std::max(rf[z], rfz[(size_t)z * kP + q]);
      rb[z] = std::max(rb[z], rbz[(size_t)z * kP + q]);
    }
  int64_t pvf = 0, pvb = 0;
  std::vector<int64_t> offf(Z), offb(Z);
  for (int64_t z = 0; z < Z; ++z) {
    offf[z] = pvf;
    offb[z] = pvb;
    pvf += plane_block(rf[z]);
    pvb += plane_block(rb[z]);
  }
  // every element is written below (zeros for taps outside the crop), so the
  // tables are not cleared first; pages are first touched by the fill threads
  std::unique_ptr<float[]> tf(new float[(size_t)pvf * V]), tb(new float[(size_t)pvb * V]);
  std::vector<int> kidx(kP, -1);
  for (int k = 0; k < V; ++k) kidx[uv[2 * k] + kN * uv[2 * k + 1]] = k;
  // forward tables, one task per (plane, view row v, lenslet row d2): the
  // task writes rows d2 of the views (u, v) (K x 225 contiguous floats each)
  // and reads the 225 raw rows y = v + 15 (d2 + R) + c0 - s2 of the planes
  // (s1, s2) that feed them; phase (s1, s2) at lenslet offset d1 is pixel
  // x = u + 15 (d1 + R) + c0 - s1 of its row
  const int NR = 2 * R + 1;
  parallel_for(Z * kN * NR, [&](int64_t q) {
    const int64_t z = q / (kN * NR);
    const int v = (int)(q / NR % kN), d2 = (int)(q % NR) - R;
    const int r = rf[z], K = 2 * r + 1;
    if (d2 < -r || d2 > r) return;
    const double hz = hs[z];
    const float* rowp[kN][kN];  // [s2][s1]; null for rows outside the crop
    for (int s2 = 0; s2 < kN; ++s2) {
      const int64_t y = v + (int64_t)kN * (d2 + R) + kC0 - s2;
      for (int s1 = 0; s1 < kN; ++s1)
        rowp[s2][s1] = y >= 0 && y < IMG ? plane_of(H, s1, s2, z) + (size_t)(i0 + y) * S + i0 : nullptr;
    }
    for (int u = 0; u < kN; ++u) {
      const int k = kidx[u + kN * v];
      if (k < 0) continue;
      float* F = tf.get() + (size_t)k * pvf + offf[z] + (size_t)(d2 + r) * K * kP;
      for (int d1 = -r; d1 <= r; ++d1) {
        const int64_t xb = u + (int64_t)kN * (d1 + R) + kC0;
        float* o = F + (size_t)(d1 + r) * kP;
        for (int s2 = 0; s2 < kN; ++s2)
          for (int s1 = 0; s1 < kN; ++s1) {
            const int64_t x = xb - s1;
            const float* rp = rowp[s2][s1];
            o[s2 * kN + s1] = rp && x >= 0 && x < IMG ? (float)((double)rp[x] / hz) : 0.f;
          }
      }
    }
  });
  // backward tables: the view's own Ht slab, s1 innermost (contiguous)
  parallel_for((int64_t)V * Z, [&](int64_t q) {
    const int64_t k = q / Z, z = q % Z;
    const int u = uv[2 * k], v = uv[2 * k + 1];
    const int qr = rb[z], Q = 2 * qr + 1;
    float* B = tb.get() + (size_t)k * pvb + offb[z];
    const float* P = plane_of(Ht, u, v, z);
    for (int d2 = -qr; d2 <= qr; ++d2)
      for (int s2 = 0; s2 < kN; ++s2) {
        const int64_t y = (int64_t)kN * (d2 + R) + kC0 + s2 - v;
        for (int d1 = -qr; d1 <= qr; ++d1)
          for (int s1 = 0; s1 < kN; ++s1) {
            const int64_t x = (int64_t)kN * (d1 + R) + kC0 + s1 - u;
            const bool in = x >= 0 && x < IMG && y >= 0 && y < IMG;
            B[((int64_t)(d2 + qr) * Q + (d1 + qr)) * kP + s2 * kN + s1] =
                in ? (float)((double)P[(size_t)(i0 + y) * S + (i0 + x)] / hts[z]) : 0.f;
          }
      }
  });
  return mlf_psf_create_compact(ctx, N, Z, V, uv.data(), wv.data(), rf.data(), tf.get(),
                                rb.data(), tb.get(), out);
}

mlf_status mlf_psf_info(mlf_psf psf, int64_t* nviews, int64_t* Z, int32_t* rf, int32_t* rb) {
  clear_error();
  if (!psf) return set_error(MLF_ERR_INVALID_ARGUMENT, "psf is NULL");
  if (nviews) *nviews = psf->V;
  if (Z) *Z = psf->Z;
  if (rf) std::copy(psf->rf.begin(), psf->rf.end(), rf);
  if (rb) std::copy(psf->rb.begin(), psf->rb.end(), rb);
  return MLF_OK;
}

mlf_status mlf_psf_views(mlf_psf psf, int32_t* uv, double* weight) {
  clear_error();
  if (!psf) return set_error(MLF_ERR_INVALID_ARGUMENT, "psf is NULL");
  if (uv) std::copy(psf->uv.begin(), psf->uv.end(), uv);
  if (weight) std::copy(psf->weight.begin(), psf->weight.end(), weight);
  return MLF_OK;
}

mlf_status mlf_psf_destroy(mlf_psf psf) {
  clear_error();
  if (!psf) return MLF_OK;
  std::lock_guard<std::mutex> lk(g_life_mu);
  if (psf->closed) return set_error(MLF_ERR_INVALID_ARGUMENT, "psf destroyed twice");
  psf->closed = true;
  if (psf->refs == 0) psf_free(psf);
  return MLF_OK;
}

}  // extern "C"

// ================================================================= patch

// Bracket a launch with a timing event pair (no-op unless timing is on).
static mlf_status tim_begin(mlf_patch_s* p, int cls) {
  if (!p->tim_now) return MLF_OK;
  if (p->ev_used + 2 > p->ev_pool.size()) {
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      MLF_HIP(hipEventCreate(&e));
      p->ev_pool.push_back(e);
    }
    p->ev_class.push_back(0);
  }
  p->ev_class[p->ev_used / 2] = cls;
  MLF_HIP(hipEventRecord(p->ev_pool[p->ev_used], p->psf->ctx->stream));
  return MLF_OK;
}

static mlf_status tim_end(mlf_patch_s* p) {
  if (!p->tim_now) return MLF_OK;
  MLF_HIP(hipEventRecord(p->ev_pool[p->ev_used + 1], p->psf->ctx->stream));
  p->ev_used += 2;
  return MLF_OK;
}

static mlf_status build_filter_plan(mlf_patch_s* p) {
  const int Z = p->Z, I1 = p->n1 * kN, I2 = p->n2 * kN;
  const std::vector<int> idx = filter_plane_index(Z);
  std::vector<int32_t> plane, n01, len1, n02, len2;
  std::vector<int64_t> off1, off2;
  std::vector<float> t1, t2;
  std::vector<int> seen;  // filter ids already computed
  std::vector<std::pair<int, std::vector<float>>> cache1, cache2;
  std::vector<std::pair<int, int>> n0cache;
  for (int z = 0; z < Z; ++z) {
    const int f = idx[z];
    if (f < 0) continue;
    const double om = filter_omega(Z, kN, f);
    int a0, b0;
    std::vector<float> ta, tbv;
    truncate_kernel(filter_kernel_1d(I1, om), &a0, &ta);
    truncate_kernel(filter_kernel_1d(I2, om), &b0, &tbv);
    plane.push_back(z);
    n01.push_back(a0);
    len1.push_back((int32_t)ta.size());
    off1.push_back((int64_t)t1.size());
    t1.insert(t1.end(), ta.begin(), ta.end());
    n02.push_back(b0);
    len2.push_back((int32_t)tbv.size());
    p->f_maxlen = std::max(p->f_maxlen, (int)std::max(ta.size(), tbv.size()));
    off2.push_back((int64_t)t2.size());
    t2.insert(t2.end(), tbv.begin(), tbv.end());
  }
  p->f_nplanes = (int)plane.size();
  if (p->f_nplanes == 0) return MLF_OK;
  MLF_CHECK(p->f_plane.alloc(plane.size()));
  MLF_CHECK(p->f_n0_1.alloc(plane.size()));
  MLF_CHECK(p->f_len_1.alloc(plane.size()));
  MLF_CHECK(p->f_off_1.alloc(plane.size()));
  MLF_CHECK(p->f_n0_2.alloc(plane.size()));
  MLF_CHECK(p->f_len_2.alloc(plane.size()));
  MLF_CHECK(p->f_off_2.alloc(plane.size()));
  MLF_CHECK(p->f_taps_1.alloc(t1.size()));
  MLF_CHECK(p->f_taps_2.alloc(t2.size()));
  MLF_HIP(hipMemcpy(p->f_plane.ptr, plane.data(), p->f_plane.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_n0_1.ptr, n01.data(), p->f_n0_1.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_len_1.ptr, len1.data(), p->f_len_1.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_off_1.ptr, off1.data(), p->f_off_1.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_n0_2.ptr, n02.data(), p->f_n0_2.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_len_2.ptr, len2.data(), p->f_len_2.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_off_2.ptr, off2.data(), p->f_off_2.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_taps_1.ptr, t1.data(), p->f_taps_1.bytes(), hipMemcpyHostToDevice));
  MLF_HIP(hipMemcpy(p->f_taps_2.ptr, t2.data(), p->f_taps_2.bytes(), hipMemcpyHostToDevice));
  p->nblk = filter_partial_blocks(p->n1, p->n2);
  MLF_CHECK(p->pb.alloc((size_t)p->nblk * p->f_nplanes));
  MLF_CHECK(p->pa.alloc((size_t)p->nblk * p->f_nplanes));
  // same row layout as the volume (rows of kXRow floats; the pad is never read)
  MLF_CHECK(p->tmp.alloc((size_t)p->f_nplanes * p->n1 * p->n2 * kXRow));
  return MLF_OK;
}

static FilterPlan filter_plan_of(mlf_patch_s* p) {
  FilterPlan fp;
  fp.nplanes = p->f_nplanes;
  fp.plane = p->f_plane.ptr;
  fp.n0_1 = p->f_n0_1.ptr;
  fp.len_1 = p->f_len_1.ptr;
  fp.off_1 = p->f_off_1.ptr;
  fp.taps_1 = p->f_taps_1.ptr;
  fp.n0_2 = p->f_n0_2.ptr;
  fp.len_2 = p->f_len_2.ptr;
  fp.off_2 = p->f_off_2.ptr;
  fp.taps_2 = p->f_taps_2.ptr;
  return fp;
}


static FwdGeo fwd_geo(mlf_patch_s* p, int k) {
  mlf_psf_s* ps = p->psf;
  const PrecPack& pk = ps->pack[(int)p->kp];
  return FwdGeo{p->d_slot.ptr, ps->d_rf.ptr, pk.d_woff.ptr, pk.d_nchunk.ptr, p->d_fsched.ptr,
                p->d_counts.ptr + 0,
                pk.d_finv.ptr ? pk.d_finv.ptr + (size_t)k * ps->Z : nullptr};
}

static BwdGeo bwd_geo(mlf_patch_s* p, int k) {
  mlf_psf_s* ps = p->psf;
  const PrecPack& pk = ps->pack[(int)p->kp];
  return BwdGeo{p->d_slot.ptr, ps->d_rb.ptr, pk.d_boff.ptr, pk.d_epad.ptr, p->d_bsched.ptr,
                p->d_counts.ptr + 1, k, bwd_groups_per_entry(p->bwd_nt, p->bwd_grp)};
}

// Forward projection of view k, then the ratio (LF != null) and/or y.
static mlf_status patch_forward_ratio(mlf_patch_s* p, int k, const float* LF, float* y_out) {
  mlf_psf_s* ps = p->psf;
  hipStream_t st = ps->ctx->stream;
  const int u = ps->uv[2 * k], v = ps->uv[2 * k + 1];
  if (p->reference) {
    const PlaneTables tf{ps->d_rf.ptr, ps->d_offf.ptr};
    MLF_CHECK(tim_begin(p, 0));
    MLF_LAUNCH(launch_forward_fp32(p->X.ptr, ps->fwd.ptr + (size_t)k * ps->per_view_f, tf, p->n1,
                                   p->n2, p->Z, p->part.ptr, st));
    MLF_CHECK(tim_end(p));
    MLF_CHECK(tim_begin(p, 1));
    MLF_LAUNCH(launch_ratio(p->part.ptr, p->Z, p->n1, p->n2, LF, u, v, p->R.ptr, y_out, st));
    MLF_CHECK(tim_end(p));
    return MLF_OK;
  }
  const PrecPack& pk = ps->pack[(int)p->kp];
  const FwdGeo g = fwd_geo(p, k);
  const void* Wk = pk.fwd.ptr + (size_t)k * pk.pvf * elem_bytes(p->kp);
  const RatioGeo rg{p->d_rsched.ptr, p->d_counts.ptr + 2};
  const int use_lf = LF != nullptr;
  MLF_CHECK(tim_begin(p, 0));
  if (p->tpz) {
    MLF_LAUNCH(launch_fwd_tpz(p->kp, p->fwd_mt, Wk, g, p->nblk_f, p->smem_f, st));
  } else {
    MLF_LAUNCH(launch_fwd_mfma(p->kp, Wk, g, p->nblk_f, p->smem_f, p->fwd_nbuf, st));
  }
  MLF_CHECK(tim_end(p));
  MLF_CHECK(tim_begin(p, 1));
  if (p->tpz) {
    MLF_LAUNCH(launch_ratio_gather_tpz(g, rg, p->nblk_r, p->Z, p->cw, u, v, use_lf, y_out, st));
  } else {
    MLF_LAUNCH(launch_ratio_gather(g, rg, p->nblk_r, p->Z, p->bm2, u, v, use_lf, y_out, st));
  }
  MLF_CHECK(tim_end(p));
  return MLF_OK;
}

static mlf_status patch_forward(mlf_patch_s* p, int k) {
  p->tim_now = p->timing;
  return patch_forward_ratio(p, k, nullptr, p->y.ptr);
}

static mlf_status patch_view_step(mlf_patch_s* p, int k) {
  mlf_psf_s* ps = p->psf;
  hipStream_t st = ps->ctx->stream;
  p->tim_now = p->timing && p->tim_seq++ % p->timing_stride == 0;
  MLF_CHECK(patch_forward_ratio(p, k, p->LF.ptr, nullptr));
  MLF_CHECK(tim_begin(p, 2));
  if (p->reference) {
    const PlaneTables tb{ps->d_rb.ptr, ps->d_offb.ptr};
    MLF_LAUNCH(launch_backward_update_fp32(p->X.ptr, ps->bwd.ptr + (size_t)k * ps->per_view_b, tb,
                                           p->R.ptr, p->n1, p->n2, p->Z, (float)ps->weight[k], st));
  } else {
    const PrecPack& pk = ps->pack[(int)p->kp];
    MLF_LAUNCH(launch_bwd_mfma(p->kp, p->bwd_nt, p->bwd_occ, p->Xlo.ptr != nullptr,
                               pk.bwd.ptr + (size_t)k * pk.pvb * elem_bytes(p->kp), bwd_geo(p, k), p->nblk_b,
                               p->smem_b, (float)ps->weight[k], st));
  }
  MLF_CHECK(tim_end(p));
  return MLF_OK;
}

// Forward partial-window elements of each plane for an n1 x n2 grid.
static int64_t partial_elems(const mlf_psf_s* ps, bool tpz, int cw, int bm2, int n1, int n2,
                             std::vector<int64_t>* poff) {
  const int colw = tpz ? cw : bm2;  // lenslets per block along axis 1
  const int nb1 = (n1 + 15) / 16, nb2 = (n2 + colw - 1) / colw;
  int64_t po = 0;
  for (int z = 0; z < ps->Z; ++z) {
    const int r = ps->rf[z];
    if (poff) (*poff)[z] = po;
    po += tpz ? (int64_t)nb1 * nb2 * 16 * (cw + 2 * r) : (int64_t)nb1 * nb2 * (16 + 2 * r) * (bm2 + 2 * r);
  }
  return po;
}

void mlf::xcd_interleave(std::vector<int32_t>& ent, int width, const std::vector<int>& group) {
  const size_t n = ent.size() / width;
  std::vector<std::vector<size_t>> lists(kXcds);
  std::vector<int> rank_of;  // group id -> rank, by first appearance
  std::vector<int> seen;
  for (size_t e = 0; e < n; ++e) {
    const int gid = group[e];
    if (gid >= (int)rank_of.size()) rank_of.resize(gid + 1, -1);
    if (rank_of[gid] < 0) rank_of[gid] = (int)seen.size(), seen.push_back(gid);
    const int g = rank_of[gid];
    const int x = ((g / kXcds) % 2 == 0) ? g % kXcds : kXcds - 1 - g % kXcds;
    lists[x].push_back(e);
  }
  std::vector<int32_t> out;
  out.reserve(ent.size());
  for (size_t k = 0;; ++k) {
    bool any = false;
    for (int x = 0; x < kXcds; ++x)
      if (k < lists[x].size()) {
        any = true;
        out.insert(out.end(), ent.begin() + lists[x][k] * width, ent.begin() + (lists[x][k] + 1) * width);
      }
    if (!any) break;
  }
  ent.swap(out);
}

// Prefix each `width`-int entry of `src` with a slot index.
static void append_entries(const std::vector<int32_t>& src, int width, int slot,
                           std::vector<int32_t>* dst) {
  for (size_t i = 0; i + width <= src.size(); i += width) {
    dst->push_back(slot);
    dst->insert(dst->end(), src.begin() + i, src.begin() + i + width);
  }
}

// Schedules (largest planes first), partial buffers, packed tables.
mlf_status mlf::patch_setup_tiles(mlf_patch_s* p) {
  mlf_psf_s* ps = p->psf;
  std::lock_guard<std::mutex> lk(ps->mu);
  static std::once_flag attr_once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(attr_once, [] { attr_err = mfma_kernel_attributes(); });
  if (attr_err != hipSuccess)
    return set_error(MLF_ERR_DEVICE, "kernel attributes: %s", hipGetErrorString(attr_err));
  MLF_CHECK(psf_ensure_pack(ps, p->kp));
  MLF_CHECK(psf_ensure_norm(ps, p->kp, p->n1, p->n2, &p->norm));
  const int Z = p->Z;
  p->tpz = ps->pack[(int)p->kp].tpz;
  p->bm2 = fwd_bm2(p->kp);
  p->cw = 4 * p->fwd_mt;
  const int colw = p->tpz ? p->cw : p->bm2;  // lenslets per block along axis 1
  const int nb1 = (p->n1 + 15) / 16, nb2 = (p->n2 + colw - 1) / colw;
  std::vector<int> order(Z);
  for (int z = 0; z < Z; ++z) order[z] = z;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ps->rf[a] > ps->rf[b]; });
  std::vector<int64_t> poff(Z);
  const int64_t po = partial_elems(ps, p->tpz, p->cw, p->bm2, p->n1, p->n2, &poff);
  int rfmax = 0, rbmax = 0, epmax = 0;
  for (int z = 0; z < Z; ++z) {
    rfmax = std::max(rfmax, ps->rf[z]);
    rbmax = std::max(rbmax, ps->rb[z]);
    epmax = std::max(epmax, ps->pack[(int)p->kp].epad[z]);
  }
  p->fent.clear();
  for (int z : order)
    for (int b2 = 0; b2 < nb2; ++b2)
      for (int b1 = 0; b1 < nb1; ++b1) {
        p->fent.push_back(z);
        p->fent.push_back(b1);
        p->fent.push_back(b2);
      }
  p->nblk_f = (int)(p->fent.size() / 3);
  {  // one plane per XCD at a time: its blocks share the plane's taps and halo rows
    std::vector<int> grp(p->nblk_f);
    for (int e = 0; e < p->nblk_f; ++e) grp[e] = p->fent[3 * e];
    xcd_interleave(p->fent, 3, grp);
  }
  std::vector<int> border(Z);
  for (int z = 0; z < Z; ++z) border[z] = z;
  std::stable_sort(border.begin(), border.end(), [&](int a, int b) { return ps->rb[a] > ps->rb[b]; });
  const int mb1 = (p->n1 + 15) / 16, mb2 = (p->n2 + 15) / 16;
  const int ngrp = kPhiPadBwd / (16 * p->bwd_nt);
  const int gpe = bwd_groups_per_entry(p->bwd_nt, p->bwd_grp);  // groups per entry share its window
  p->bent.clear();
  for (int z : border)
    for (int b2 = 0; b2 < mb2; ++b2)
      for (int b1 = 0; b1 < mb1; ++b1)
        for (int gq = 0; gq < ngrp; gq += gpe) {
          p->bent.push_back(z);
          p->bent.push_back(b1);
          p->bent.push_back(b2);
          p->bent.push_back(gq);
        }
  p->nblk_b = (int)(p->bent.size() / 4);
  {  // one (plane, phase group) per XCD at a time: shared tap table
    std::vector<int> grp(p->nblk_b);
    for (int e = 0; e < p->nblk_b; ++e) grp[e] = p->bent[4 * e] * ngrp + p->bent[4 * e + 3];
    xcd_interleave(p->bent, 4, grp);
  }
  p->nblk_r = (p->n1 * p->n2 + kRatioLens - 1) / kRatioLens;
  std::vector<int32_t> fs, bs, rs;
  append_entries(p->fent, 3, 0, &fs);
  append_entries(p->bent, 4, 0, &bs);
  for (int i = 0; i < p->nblk_r; ++i) {
    rs.push_back(0);
    rs.push_back(kRatioLens * i);
  }
  MLF_CHECK(upload_vec(p->d_poff, poff));
  MLF_CHECK(upload_vec(p->d_fsched, fs));
  MLF_CHECK(upload_vec(p->d_bsched, bs));
  MLF_CHECK(upload_vec(p->d_rsched, rs));
  MLF_CHECK(upload_vec(p->d_counts, std::vector<int32_t>{p->nblk_f, p->nblk_b, p->nblk_r}));
  MLF_CHECK(p->partial.alloc((size_t)po));
  if (p->tpz && p->kp != Precision::kF32)
    MLF_CHECK(p->Xb.alloc((size_t)p->n1 * p->n2 * Z * kShadowRow));
  // BF16: the view loop keeps the volume split into the shadow (upper half)
  // and a lower plane, so the backprojection writes 4 bytes per voxel instead
  // of FP32 X plus the shadow
  if (p->tpz && p->kp == Precision::kBF16)
    MLF_CHECK(p->Xlo.alloc((size_t)p->n1 * p->n2 * Z * kShadowRow));
  SlotDesc& sd = p->slot;
  sd.X = p->X.ptr;
  sd.Xb = reinterpret_cast<__bf16*>(p->Xb.ptr);
  sd.Xlo = p->Xlo.ptr;
  sd.LF = p->LF.ptr;
  sd.R = p->R.ptr;
  sd.partial = p->partial.ptr;
  sd.poff = p->d_poff.ptr;
  sd.Tn = p->norm->T.ptr;
  sd.pvt = p->norm->pvt;
  sd.toffT = p->norm->d_toff.ptr;
  sd.C1 = p->norm->d_C1.ptr;
  sd.n1 = p->n1;
  sd.n2 = p->n2;
  sd.xs = p->xs;
  sd.xs_inv = 1.f / p->xs;
  MLF_CHECK(upload_vec(p->d_slot, std::vector<SlotDesc>{sd}));
  p->smem_f = p->tpz ? tpz_smem_bytes(p->kp, rfmax, p->fwd_mt) : fwd_smem_bytes(p->kp, rfmax, p->bm2);
  p->fwd_nbuf = fwd_nbuf(p->kp, rfmax, p->bm2);
  p->smem_b = bwd_smem_bytes(p->kp, rbmax, epmax, p->bwd_nt, p->bwd_occ);
  if (p->smem_f > 160 * 1024 || p->smem_b > 160 * 1024)
    return set_error(MLF_ERR_UNSUPPORTED,
                     "PSF radius too large for the LDS tiles (rf_max=%d, rb_max=%d)", rfmax, rbmax);
  return MLF_OK;
}

// Re-derive the 16-bit shadow (BF16 Toeplitz: the split form) after X changed
// outside the backprojection.
mlf_status mlf::patch_refresh_shadow(mlf_patch_s* p) {
  if (!p->Xb.ptr) return MLF_OK;
  const size_t nl = (size_t)p->n1 * p->n2 * p->Z;
  if (p->Xlo.ptr)
    MLF_LAUNCH(launch_make_split(p->X.ptr, p->Xb.ptr, p->Xlo.ptr, nl, p->psf->ctx->stream));
  else
    MLF_LAUNCH(launch_make_shadow(p->kp, p->X.ptr, p->Xb.ptr, nl, p->slot.xs, p->psf->ctx->stream));
  return MLF_OK;
}

// Rebuild the FP32 volume from the split form after view steps (exact).
mlf_status mlf::patch_merge_split(mlf_patch_s* p) {
  if (!p->Xlo.ptr) return MLF_OK;
  MLF_LAUNCH(launch_merge_split(p->Xb.ptr, p->Xlo.ptr, p->X.ptr, (size_t)p->n1 * p->n2 * p->Z,
                                p->psf->ctx->stream));
  return MLF_OK;
}

mlf_status mlf::patch_filter(mlf_patch_s* p) {
  if (p->f_nplanes == 0) return MLF_OK;
  hipStream_t st = p->psf->ctx->stream;
  const FilterPlan fp = filter_plan_of(p);
  p->tim_now = p->timing;
  MLF_CHECK(tim_begin(p, 3));
  MLF_LAUNCH(launch_filter_pass1(p->X.ptr, p->tmp.ptr, fp, p->n1, p->n2, p->pb.ptr, p->f_maxlen, st));
  MLF_LAUNCH(launch_filter_pass2(p->tmp.ptr, p->X.ptr, fp, p->n1, p->n2, p->pa.ptr, p->f_maxlen, st));
  MLF_LAUNCH(launch_filter_scale(p->X.ptr, fp, p->n1, p->n2, p->pb.ptr, p->pa.ptr, st));
  MLF_CHECK(patch_refresh_shadow(p));
  MLF_CHECK(tim_end(p));
  return MLF_OK;
}

size_t mlf::patch_device_bytes(const mlf_patch_s* p) {
  size_t b = p->X.bytes() + p->LF.bytes() + p->part.bytes() + p->R.bytes() + p->y.bytes() +
             p->tmp.bytes() + p->pb.bytes() + p->pa.bytes() + p->partial.bytes() + p->Xb.bytes() +
             p->Xlo.bytes();
  b += p->d_fsched.bytes() + p->d_bsched.bytes() + p->d_rsched.bytes() + p->f_taps_1.bytes() +
       p->f_taps_2.bytes();
  return b;
}

// voxels of the patch's volume (the MATLAB array's element count)
static size_t patch_numel(const mlf_patch_s* p) { return (size_t)p->n1 * p->n2 * p->Z * kP; }

size_t mlf::patch_bytes_estimate(mlf_psf_s* ps, mlf_precision prec, int n1, int n2) {
  const Precision kp = kernel_precision(prec);
  int rfmax = 0;
  for (int z = 0; z < ps->Z; ++z) rfmax = std::max(rfmax, ps->rf[z]);
  const bool tpz = rfmax <= tpz_max_radius();
  const int cw = 4 * (kp == Precision::kF32 ? kTpzMTF32 : kTpzMTBF16);
  const size_t n = (size_t)n1 * n2, Z = (size_t)ps->Z;
  size_t b = 4 * n * Z * kXRow * 2;  // volume + sample_filter temporary
  b += 4 * n * kP + 4 * n * Z + 8 * n;
  b += 4 * (size_t)partial_elems(ps, tpz, cw, fwd_bm2(kp), n1, n2, nullptr);
  if (kp != Precision::kF32 && tpz) b += 2 * n * Z * kShadowRow;
  if (kp == Precision::kBF16 && tpz) b += 2 * n * Z * kShadowRow;  // lower plane of the split volume
  // entry lists and filter taps: < 1 MB
  return b + (1u << 20);
}

extern "C" {

mlf_status mlf_patch_create(mlf_psf psf, int64_t n1, int64_t n2, mlf_precision prec,
                            mlf_patch* out) {
  clear_error();
  if (!psf || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (n1 < 1 || n2 < 1 || n1 > 4096 || n2 > 4096)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad lenslet grid %lld x %lld", (long long)n1,
                     (long long)n2);
  if (prec != MLF_PRECISION_FP32 && prec != MLF_PRECISION_BF16 && prec != MLF_PRECISION_FP16)
    return set_error(MLF_ERR_UNSUPPORTED, "precision %d not supported", (int)prec);
  // kernels address one plane of the volume with 32-bit byte offsets
  if ((double)n1 * n2 * kXRow * sizeof(float) >= 2147483648.0)
    return set_error(MLF_ERR_UNSUPPORTED, "lenslet grid %lld x %lld: a plane exceeds 2 GiB",
                     (long long)n1, (long long)n2);
  MLF_CHECK(ctx_activate(psf->ctx));
  auto p = std::make_unique<mlf_patch_s>();
  p->psf = psf;
  {
    std::lock_guard<std::mutex> lk(g_life_mu);
    if (psf->closed) return set_error(MLF_ERR_INVALID_ARGUMENT, "psf already destroyed");
    ++psf->refs;
  }
  struct RefUndo {
    mlf_patch_s* p;
    ~RefUndo() {
      if (p) {
        std::lock_guard<std::mutex> lk(g_life_mu);
        --p->psf->refs;
      }
    }
  } undo{p.get()};
  p->n1 = (int)n1;
  p->n2 = (int)n2;
  p->Z = psf->Z;
  p->prec = prec;
  const size_t n = (size_t)n1 * n2;
  // volume rows are kXRow floats; the pad is zeroed here and never written again
  MLF_CHECK(p->X.alloc(n * p->Z * kXRow));
  MLF_HIP(hipMemsetAsync(p->X.ptr, 0, p->X.bytes(), psf->ctx->stream));
  MLF_HIP(hipStreamSynchronize(psf->ctx->stream));
  MLF_CHECK(p->LF.alloc(n * kP));
  MLF_CHECK(p->part.alloc(n * p->Z));
  MLF_CHECK(p->R.alloc(n));
  MLF_CHECK(p->y.alloc(n));
  MLF_CHECK(p->dsum.alloc(1));
  MLF_CHECK(build_filter_plan(p.get()));
  const char* ref = std::getenv("MLF_REFERENCE_KERNELS");
  p->reference = ref && ref[0] == '1' && prec == MLF_PRECISION_FP32;
  p->kp = kernel_precision(prec);
  if (!p->reference) {
    const TileShape t0 = tile_candidates(p->kp)[0];
    p->fwd_mt = t0.mt;
    p->bwd_nt = t0.nt;
    p->bwd_grp = t0.grp;
    p->bwd_occ = t0.occ;
    TileShape t;
    if (tune_cache_get(tune_key(p.get(), "patch"), p->kp, &t)) {
      p->fwd_mt = t.mt;
      p->bwd_nt = t.nt;
      p->bwd_grp = t.grp;
      p->bwd_occ = t.occ;
    }
    MLF_CHECK(patch_setup_tiles(p.get()));
  }
  MLF_HIP(hipEventCreate(&p->ev0));
  MLF_HIP(hipEventCreate(&p->ev1));
  undo.p = nullptr;
  *out = p.release();
  return MLF_OK;
}

mlf_status mlf_patch_destroy(mlf_patch p) {
  clear_error();
  if (!p) return MLF_OK;
  std::lock_guard<std::mutex> lk(g_life_mu);
  mlf_psf psf = p->psf;
  (void)hipSetDevice(psf->ctx->device);
  (void)batch_forget_patch(p);
  (void)hipStreamSynchronize(psf->ctx->stream);
  delete p;
  if (--psf->refs == 0 && psf->closed) psf_free(psf);
  return MLF_OK;
}

mlf_status mlf_patch_set_image(mlf_patch p, const float* lfimg) {
  clear_error();
  if (!p || !lfimg) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  MLF_HIP(hipMemcpyAsync(p->LF.ptr, lfimg, p->LF.bytes(), hipMemcpyHostToDevice,
                         p->psf->ctx->stream));
  MLF_HIP(hipStreamSynchronize(p->psf->ctx->stream));
  p->image_set = true;
  return MLF_OK;
}

// FP16 shadow scale: X * xs ~ 1 for a volume of mean `mean` (kernels in
// flight read the slot descriptor, so the stream is synchronised first).
static mlf_status patch_set_scale(mlf_patch_s* p, double mean) {
  if (p->kp != Precision::kF16 || p->reference) return MLF_OK;
  const float xs = (mean > 0.0 && std::isfinite(mean)) ? (float)(1.0 / mean) : 1.f;
  if (xs == p->xs) return MLF_OK;
  MLF_HIP(hipStreamSynchronize(p->psf->ctx->stream));
  p->xs = xs;
  p->slot.xs = xs;
  p->slot.xs_inv = 1.f / xs;
  MLF_CHECK(upload_vec(p->d_slot, std::vector<SlotDesc>{p->slot}));
  if (p->batch) batch_mark_dirty(p->batch);
  return MLF_OK;
}

mlf_status mlf_patch_init_volume(mlf_patch p) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->image_set) return set_error(MLF_ERR_INVALID_ARGUMENT, "image not set");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  // recon:72-73: X = ones ./ numel .* sum(LFIMG), FP64 image sum on the device.
  // Asynchronous on the context stream.
  const double numel = (double)p->n1 * kN * p->n2 * kN * p->Z;
  MLF_LAUNCH(launch_init_volume(p->LF.ptr, p->LF.count, p->dsum.ptr, p->X.ptr, p->X.count,
                                1.0 / numel, p->psf->ctx->stream));
  if (p->kp == Precision::kF16) {
    double sum = 0.0;
    MLF_HIP(hipMemcpyAsync(&sum, p->dsum.ptr, sizeof(double), hipMemcpyDeviceToHost, p->psf->ctx->stream));
    MLF_HIP(hipStreamSynchronize(p->psf->ctx->stream));
    MLF_CHECK(patch_set_scale(p, sum / numel));
  }
  MLF_CHECK(patch_refresh_shadow(p));
  p->volume_set = true;
  return MLF_OK;
}

mlf_status mlf_patch_set_volume(mlf_patch p, const float* X) {
  clear_error();
  if (!p || !X) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  const size_t numel = patch_numel(p);
  DevBuf<float> tmp;
  MLF_CHECK(tmp.alloc(numel));
  MLF_HIP(hipMemcpyAsync(tmp.ptr, X, tmp.bytes(), hipMemcpyHostToDevice, st));
  MLF_LAUNCH(launch_to_blocked(tmp.ptr, p->X.ptr, p->n1, p->n2, p->Z, st));
  if (p->kp == Precision::kF16) {
    double sum = 0.0;
    for (size_t i = 0; i < numel; ++i) sum += X[i];
    MLF_CHECK(patch_set_scale(p, sum / (double)numel));
  }
  MLF_CHECK(patch_refresh_shadow(p));
  MLF_HIP(hipStreamSynchronize(st));
  p->volume_set = true;
  return MLF_OK;
}

static mlf_status get_volume_impl(mlf_patch p, void* out, bool f64) {
  clear_error();
  if (!p || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  if (f64) {
    DevBuf<double> tmp;
    MLF_CHECK(tmp.alloc(patch_numel(p)));
    MLF_LAUNCH(launch_from_blocked_f64(p->X.ptr, tmp.ptr, p->n1, p->n2, p->Z, st));
    MLF_HIP(hipMemcpyAsync(out, tmp.ptr, tmp.bytes(), hipMemcpyDeviceToHost, st));
    MLF_HIP(hipStreamSynchronize(st));
  } else {
    DevBuf<float> tmp;
    MLF_CHECK(tmp.alloc(patch_numel(p)));
    MLF_LAUNCH(launch_from_blocked_f32(p->X.ptr, tmp.ptr, p->n1, p->n2, p->Z, st));
    MLF_HIP(hipMemcpyAsync(out, tmp.ptr, tmp.bytes(), hipMemcpyDeviceToHost, st));
    MLF_HIP(hipStreamSynchronize(st));
  }
  return MLF_OK;
}

mlf_status mlf_patch_get_volume(mlf_patch p, float* X) { return get_volume_impl(p, X, false); }
mlf_status mlf_patch_get_volume_f64(mlf_patch p, double* X) { return get_volume_impl(p, X, true); }

mlf_status mlf_patch_run(mlf_patch p, int32_t iterations, int32_t views_limit, int32_t filter) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->image_set || !p->volume_set)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "image and volume must be set before running");
  if (iterations < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "negative iteration count");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  const int nv = (views_limit <= 0 || views_limit > p->psf->V) ? p->psf->V : views_limit;
  MLF_HIP(hipEventRecord(p->ev0, st));
  for (int it = 0; it < iterations; ++it) {
    for (int k = 0; k < nv; ++k) MLF_CHECK(patch_view_step(p, k));
    MLF_CHECK(patch_merge_split(p));
    if (filter) MLF_CHECK(patch_filter(p));
  }
  MLF_HIP(hipEventRecord(p->ev1, st));
  p->timed = true;
  return MLF_OK;
}

mlf_status mlf_patch_last_run_ms(mlf_patch p, double* ms) {
  clear_error();
  if (!p || !ms) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->timed) return set_error(MLF_ERR_INVALID_ARGUMENT, "no run recorded");
  MLF_CHECK(ctx_activate(p->psf->ctx));
  MLF_HIP(hipEventSynchronize(p->ev1));
  float f = 0.f;
  MLF_HIP(hipEventElapsedTime(&f, p->ev0, p->ev1));
  *ms = f;
  return MLF_OK;
}

mlf_status mlf_patch_set_timing(mlf_patch p, int32_t enable) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (enable < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "negative timing stride");
  std::lock_guard<std::mutex> lk(p->mu);
  p->timing = enable != 0;
  p->timing_stride = enable > 1 ? enable : 1;
  p->tim_seq = 0;
  p->tim_now = false;
  p->ev_used = 0;
  return MLF_OK;
}

mlf_status mlf_patch_kernel_times(mlf_patch p, double* total_ms, int64_t* launches) {
  clear_error();
  if (!p || !total_ms || !launches) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  for (int c = 0; c < MLF_KERNEL_CLASSES; ++c) {
    total_ms[c] = 0.0;
    launches[c] = 0;
  }
  if (p->ev_used) MLF_HIP(hipEventSynchronize(p->ev_pool[p->ev_used - 1]));
  for (size_t k = 0; k + 1 < p->ev_used; k += 2) {
    float ms = 0.f;
    MLF_HIP(hipEventElapsedTime(&ms, p->ev_pool[k], p->ev_pool[k + 1]));
    const int c = p->ev_class[k / 2];
    total_ms[c] += ms;
    launches[c] += 1;
  }
  p->ev_used = 0;
  return MLF_OK;
}

mlf_status mlf_patch_forward_view(mlf_patch p, int32_t view, float* y) {
  clear_error();
  if (!p || !y) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (view < 0 || view >= p->psf->V) return set_error(MLF_ERR_INVALID_ARGUMENT, "view out of range");
  if (!p->volume_set) return set_error(MLF_ERR_INVALID_ARGUMENT, "volume not set");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  MLF_CHECK(patch_forward(p, view));
  MLF_HIP(hipMemcpyAsync(y, p->y.ptr, p->y.bytes(), hipMemcpyDeviceToHost, st));
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

mlf_status mlf_patch_project_all(mlf_patch p, float* sensor) {
  clear_error();
  if (!p || !sensor) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->volume_set) return set_error(MLF_ERR_INVALID_ARGUMENT, "volume not set");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  DevBuf<float> img;
  MLF_CHECK(img.alloc(p->LF.count));
  MLF_HIP(hipMemsetAsync(img.ptr, 0, img.bytes(), st));
  for (int k = 0; k < p->psf->V; ++k) {
    MLF_CHECK(patch_forward(p, k));
    MLF_LAUNCH(launch_scatter_view(p->y.ptr, 1, p->n1, p->n2, p->psf->uv[2 * k],
                                   p->psf->uv[2 * k + 1], img.ptr, st));
  }
  MLF_HIP(hipMemcpyAsync(sensor, img.ptr, img.bytes(), hipMemcpyDeviceToHost, st));
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

mlf_status mlf_patch_sample_filter(mlf_patch p) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  MLF_CHECK(patch_filter(p));
  MLF_HIP(hipStreamSynchronize(p->psf->ctx->stream));
  return MLF_OK;
}

mlf_status mlf_patch_postprocess(mlf_patch p, int32_t valid_range, double* out, uint16_t* tiff) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!p->volume_set) return set_error(MLF_ERR_INVALID_ARGUMENT, "volume not set");
  const int Z = p->Z;
  // LFM_reconstruction_module.m:46 reads planes 1 and 2
  if (Z < 2)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "post-processing needs at least 2 planes, got %d", Z);
  // mesolf_set_params.m:138: a positive odd number of planes, centred at
  // ceil(Z/2); 0 keeps every plane (the module's own output, before the crop)
  if (valid_range < 0 || (valid_range > 0 && valid_range % 2 == 0))
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "valid_range must be a positive odd integer (or 0: all planes), got %d",
                     (int)valid_range);
  const int zc = (Z + 1) / 2, h = valid_range / 2;
  const int zlo = valid_range ? zc - h - 1 : 0, zhi = valid_range ? zc + h : Z;  // 0-based, [zlo, zhi)
  if (zlo < 0 || zhi > Z)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "valid_range %d selects planes %d..%d outside 1..%d",
                     (int)valid_range, zc - h, zc + h, Z);
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  const size_t plane = (size_t)p->n1 * kN * p->n2 * kN;
  DevBuf<double> coef, parts, dout;
  DevBuf<uint16_t> dtiff;
  MLF_CHECK(coef.alloc(4 * (size_t)Z + 2));
  MLF_CHECK(parts.alloc((size_t)kPostBlocks * Z));
  if (out) MLF_CHECK(dout.alloc(plane * (zhi - zlo)));
  if (tiff) MLF_CHECK(dtiff.alloc(plane * Z));
  p->tim_now = p->timing;
  MLF_CHECK(tim_begin(p, 4));
  MLF_LAUNCH(launch_postprocess(p->X.ptr, p->n1, p->n2, Z, coef.ptr, parts.ptr, zlo, zhi, dout.ptr,
                                dtiff.ptr, st));
  MLF_CHECK(tim_end(p));
  if (out) MLF_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, st));
  if (tiff) MLF_HIP(hipMemcpyAsync(tiff, dtiff.ptr, dtiff.bytes(), hipMemcpyDeviceToHost, st));
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

mlf_status mlf_patch_tile_shape(mlf_patch p, mlf_tile_shape* cur) {
  clear_error();
  if (!p || !cur) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  cur->fwd_columns_per_wave = p->fwd_mt;
  cur->bwd_phase_tiles = p->bwd_nt;
  cur->bwd_groups_per_entry = p->bwd_grp;
  cur->bwd_waves_per_simd = p->bwd_occ;
  cur->ms_per_view = 0.0;
  return MLF_OK;
}

mlf_status mlf_patch_set_tile_shape(mlf_patch p, const mlf_tile_shape* shape) {
  clear_error();
  if (!p || !shape) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (p->reference) return set_error(MLF_ERR_UNSUPPORTED, "reference kernels have one shape");
  if (p->batch) return set_error(MLF_ERR_INVALID_ARGUMENT, "detach the patch from its batch first");
  const TileShape want{shape->fwd_columns_per_wave, shape->bwd_phase_tiles, shape->bwd_groups_per_entry,
                       shape->bwd_waves_per_simd};
  bool ok = false;
  for (const auto& c : tile_candidates(p->kp)) ok = ok || c == want;
  if (!ok)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "tile shape (%d, %d, %d, %d) is not built for this precision",
                     (int)shape->fwd_columns_per_wave, (int)shape->bwd_phase_tiles,
                     (int)shape->bwd_groups_per_entry, (int)shape->bwd_waves_per_simd);
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  MLF_HIP(hipStreamSynchronize(p->psf->ctx->stream));
  const int mt = p->fwd_mt, nt = p->bwd_nt, gr = p->bwd_grp, oc = p->bwd_occ;
  p->fwd_mt = want.mt;
  p->bwd_nt = want.nt;
  p->bwd_grp = want.grp;
  p->bwd_occ = want.occ;
  if (patch_setup_tiles(p) != MLF_OK) {
    const std::string e = last_error();
    p->fwd_mt = mt;
    p->bwd_nt = nt;
    p->bwd_grp = gr;
    p->bwd_occ = oc;
    MLF_CHECK(patch_setup_tiles(p));
    return set_error(MLF_ERR_UNSUPPORTED, "%s", e.c_str());
  }
  if (p->volume_set) MLF_CHECK(patch_refresh_shadow(p));
  return MLF_OK;
}

mlf_status mlf_patch_autotune(mlf_patch p, int32_t views, mlf_tile_shape* best) {
  clear_error();
  if (!p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (p->reference) return set_error(MLF_ERR_UNSUPPORTED, "reference kernels have one shape");
  if (p->batch) return set_error(MLF_ERR_INVALID_ARGUMENT, "detach the patch from its batch first");
  if (!p->image_set || !p->volume_set)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "image and volume must be set before tuning");
  std::lock_guard<std::mutex> lk(p->mu);
  MLF_CHECK(ctx_activate(p->psf->ctx));
  hipStream_t st = p->psf->ctx->stream;
  const int nv = std::max(1, std::min<int>(views <= 0 ? 8 : views, p->psf->V));
  DevBuf<float> save;
  MLF_CHECK(save.alloc(p->X.count));
  MLF_HIP(hipMemcpyAsync(save.ptr, p->X.ptr, p->X.bytes(), hipMemcpyDeviceToDevice, st));
  struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
      if (e) (void)hipEventDestroy(e);
    }
  } e0, e1;
  MLF_HIP(hipEventCreate(&e0.e));
  MLF_HIP(hipEventCreate(&e1.e));
  double best_ms = 1e300;
  TileShape bt{p->fwd_mt, p->bwd_nt, p->bwd_grp, p->bwd_occ};
  // two rounds over the candidates, each candidate keeping its faster round:
  // shapes a few per cent apart are otherwise decided by clock and cache noise
  const auto cands = tile_candidates(p->kp);
  std::vector<double> cand_ms(cands.size(), 1e300);
  for (int round = 0; round < 2; ++round) {
    for (size_t ci = 0; ci < cands.size(); ++ci) {
      const auto& c = cands[ci];
      MLF_HIP(hipStreamSynchronize(st));
      p->fwd_mt = c.mt;
      p->bwd_nt = c.nt;
      p->bwd_grp = c.grp;
      p->bwd_occ = c.occ;
      if (patch_setup_tiles(p) != MLF_OK) continue;  // shape does not fit this PSF
      MLF_CHECK(patch_refresh_shadow(p));
      MLF_CHECK(patch_view_step(p, 0));  // warm-up
      MLF_HIP(hipEventRecord(e0.e, st));
      for (int k = 0; k < nv; ++k) MLF_CHECK(patch_view_step(p, k));
      MLF_HIP(hipEventRecord(e1.e, st));
      MLF_HIP(hipEventSynchronize(e1.e));
      float ms = 0.f;
      MLF_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
      MLF_HIP(hipMemcpyAsync(p->X.ptr, save.ptr, p->X.bytes(), hipMemcpyDeviceToDevice, st));
      cand_ms[ci] = std::min(cand_ms[ci], (double)ms / nv);
    }
  }
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    if (cand_ms[ci] < best_ms) {
      best_ms = cand_ms[ci];
      bt = cands[ci];
    }
  }
  MLF_HIP(hipStreamSynchronize(st));
  p->fwd_mt = bt.mt;
  p->bwd_nt = bt.nt;
  p->bwd_grp = bt.grp;
  p->bwd_occ = bt.occ;
  MLF_CHECK(patch_setup_tiles(p));
  MLF_CHECK(patch_refresh_shadow(p));
  tune_cache_put(tune_key(p, "patch"), bt, best_ms);
  if (best) {
    best->fwd_columns_per_wave = bt.mt;
    best->bwd_phase_tiles = bt.nt;
    best->bwd_groups_per_entry = bt.grp;
    best->bwd_waves_per_simd = bt.occ;
    best->ms_per_view = best_ms;
  }
  return MLF_OK;
}

// ============================================================ drop-ins
mlf_status mlf_recon_phase_space(const float* lfimg, int64_t I1, int64_t I2, const float* H,
                                 const float* Ht, int64_t S, int64_t N, int64_t Z,
                                 int32_t max_iter, int32_t gpu_flag, int32_t device,
                                 mlf_precision prec, double* X_out, mlf_recon_stats* stats) {
  clear_error();
  if (!lfimg || !H || !Ht || !X_out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!gpu_flag)
    return set_error(MLF_ERR_UNSUPPORTED,
                     "gpuFlag=false: the reference CPU branch cannot run as written "
                     "(undefined Xguessv at recon:112); use the GPU path");
  if (N != kN) return set_error(MLF_ERR_UNSUPPORTED, "Nnum must be 15");
  if (I1 % N != 0 || I2 % N != 0 || I1 < N || I2 < N)
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "LFIMG size %lld x %lld is not a positive multiple of Nnum", (long long)I1,
                     (long long)I2);
  if (max_iter < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "maxIter < 0");
  const auto t0 = std::chrono::steady_clock::now();
  // The reference pipeline calls recon once per patch with the same H and Ht
  // (LFM_reconstruction_module.m:37 inside mesolf.m's patch loop).  The last
  // prepared PSF stays resident with its context and is reused when device,
  // sizes and the contents of H and Ht (a 64-bit content hash of each) match;
  // MLF_RECON_PSF_CACHE=0 turns this off.
  struct PsfCache {
    std::mutex mu;
    mlf_context ctx = nullptr;
    mlf_psf psf = nullptr;
    int32_t device = -1;
    int64_t S = 0, Z = 0;
    uint64_t hH = 0, hHt = 0;
  };
  static PsfCache cache;
  const char* env = std::getenv("MLF_RECON_PSF_CACHE");
  const bool use_cache = !(env && env[0] == '0');
  const size_t nH = (size_t)S * S * N * N * Z;
  std::unique_lock<std::mutex> cache_lk(cache.mu, std::defer_lock);
  mlf_psf psf = nullptr;
  mlf_context ctx = nullptr;
  std::unique_ptr<mlf_context_s, mlf_status (*)(mlf_context)> ctx_guard(nullptr, mlf_context_destroy);
  std::unique_ptr<mlf_psf_s, mlf_status (*)(mlf_psf)> psf_guard(nullptr, mlf_psf_destroy);
  if (use_cache) {
    cache_lk.lock();  // one drop-in reconstruction at a time shares the cached PSF
    const uint64_t hH = content_hash(H, nH), hHt = content_hash(Ht, nH);
    if (!(cache.psf && cache.device == device && cache.S == S && cache.Z == Z && cache.hH == hH &&
          cache.hHt == hHt)) {
      if (cache.psf) (void)mlf_psf_destroy(cache.psf);
      if (cache.ctx) (void)mlf_context_destroy(cache.ctx);
      cache.psf = nullptr;
      cache.ctx = nullptr;
      MLF_CHECK(mlf_context_create(device, &cache.ctx));
      const mlf_status cs = mlf_psf_create_dense(cache.ctx, H, Ht, S, N, Z, &cache.psf);
      if (cs != MLF_OK) {
        const std::string e = last_error();
        (void)mlf_context_destroy(cache.ctx);
        cache.ctx = nullptr;
        cache.psf = nullptr;
        return set_error(cs, "%s", e.c_str());
      }
      cache.device = device;
      cache.S = S;
      cache.Z = Z;
      cache.hH = hH;
      cache.hHt = hHt;
    }
    ctx = cache.ctx;
    psf = cache.psf;
  } else {
    MLF_CHECK(mlf_context_create(device, &ctx));
    ctx_guard.reset(ctx);
    MLF_CHECK(mlf_psf_create_dense(ctx, H, Ht, S, N, Z, &psf));
    psf_guard.reset(psf);
  }
  mlf_patch p = nullptr;
  MLF_CHECK(mlf_patch_create(psf, I1 / N, I2 / N, prec, &p));
  std::unique_ptr<mlf_patch_s, mlf_status (*)(mlf_patch)> patch_guard(p, mlf_patch_destroy);
  MLF_CHECK(mlf_patch_set_image(p, lfimg));
  MLF_CHECK(mlf_patch_init_volume(p));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->views_per_iteration = psf->V;
  }
  // relative change per iteration, FP64 on the device (stats only)
  constexpr int kNormBlocks = 1024;
  DevBuf<float> prev;
  DevBuf<double> parts;
  if (stats) {
    MLF_CHECK(prev.alloc(p->X.count));
    MLF_CHECK(parts.alloc(2 * kNormBlocks));
  }
  std::vector<double> hparts(2 * kNormBlocks);
  hipStream_t st = psf->ctx->stream;
  for (int it = 0; it < max_iter; ++it) {
    if (stats)
      MLF_HIP(hipMemcpyAsync(prev.ptr, p->X.ptr, p->X.bytes(), hipMemcpyDeviceToDevice, st));
    MLF_CHECK(mlf_patch_run(p, 1, 0, 1));
    double ms = 0.0;
    MLF_CHECK(mlf_patch_last_run_ms(p, &ms));
    if (stats) {
      MLF_LAUNCH(launch_diff_norms(p->X.ptr, prev.ptr, p->X.count, parts.ptr, kNormBlocks, st));
      MLF_HIP(hipMemcpyAsync(hparts.data(), parts.ptr, parts.bytes(), hipMemcpyDeviceToHost, st));
      MLF_HIP(hipStreamSynchronize(st));
      double dd = 0.0, bb = 0.0;
      for (int k = 0; k < kNormBlocks; ++k) {
        dd += hparts[2 * k];
        bb += hparts[2 * k + 1];
      }
      if (it < 64) {
        stats->iteration_seconds[it] = ms * 1e-3;
        stats->relative_change[it] = bb > 0.0 ? std::sqrt(dd / bb) : 0.0;
      }
      stats->iterations = it + 1;
    }
  }
  MLF_CHECK(mlf_patch_get_volume_f64(p, X_out));
  if (stats)
    stats->total_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return MLF_OK;
}

mlf_status mlf_lfm_convolution_vec_device(mlf_context ctx, const float* d_bitmap, int64_t width,
                                          int64_t height, int64_t n, const float* d_kernels,
                                          int64_t kw, int64_t nphase, float* d_out) {
  clear_error();
  if (!ctx || !d_bitmap || !d_kernels || !d_out)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width < 1 || height < 1 || n < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bitmap dims must be >= 1");
  if (nphase != kN) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernels must be [kw,kw,15,15]");
  if (kw < 1 || kw > 2048) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernel width %lld out of [1,2048]", (long long)kw);
  if (width > (1 << 30) / 1 || height > (1 << 30) || width * height > (int64_t)1 << 40)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bitmap too large");
  if (n > 65535) return set_error(MLF_ERR_INVALID_ARGUMENT, "batch too large (n > 65535)");
  MLF_CHECK(ctx_activate(ctx));
  if (kw % 2 == 1 && kw <= kMexMfmaMaxKw) {
    // odd kernels: the MFMA form; its operand scratch lives in the context
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t need = mex_kq_floats((int)kw);
    if (ctx->mex_kq.count < need) {
      MLF_HIP(hipStreamSynchronize(ctx->stream));  // earlier calls may still read it
      MLF_CHECK(ctx->mex_kq.alloc(need));
    }
    MLF_LAUNCH(launch_mex_mfma(d_bitmap, width, height, n, d_kernels, (int)kw, ctx->mex_kq.ptr,
                               d_out, ctx->stream));
    return MLF_OK;
  }
  MLF_LAUNCH(launch_mex_conv(d_bitmap, width, height, n, d_kernels, (int)kw, d_out, ctx->stream));
  return MLF_OK;
}

mlf_status mlf_lfm_convolution_vec(const float* bitmap, int64_t width, int64_t height, int64_t n,
                                   const float* kernels, int64_t kw, int64_t nphase, float* out,
                                   int32_t device) {
  clear_error();
  if (!bitmap || !kernels || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width < 1 || height < 1 || n < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bitmap dims must be >= 1");
  if (nphase != kN) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernels must be [kw,kw,15,15]");
  if (kw < 1 || kw > 2048) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernel width out of [1,2048]");
  mlf_context ctx = nullptr;
  MLF_CHECK(mlf_context_create(device, &ctx));
  std::unique_ptr<mlf_context_s, mlf_status (*)(mlf_context)> guard(ctx, mlf_context_destroy);
  const size_t nb = (size_t)width * height * n;
  const size_t nk = (size_t)kw * kw * kP;
  DevBuf<float> db, dk, dout;
  MLF_CHECK(db.alloc(nb));
  MLF_CHECK(dk.alloc(nk));
  MLF_CHECK(dout.alloc(nb));
  MLF_HIP(hipMemcpyAsync(db.ptr, bitmap, db.bytes(), hipMemcpyHostToDevice, ctx->stream));
  MLF_HIP(hipMemcpyAsync(dk.ptr, kernels, dk.bytes(), hipMemcpyHostToDevice, ctx->stream));
  MLF_CHECK(mlf_lfm_convolution_vec_device(ctx, db.ptr, width, height, n, dk.ptr, kw, nphase,
                                           dout.ptr));
  MLF_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  MLF_HIP(hipStreamSynchronize(ctx->stream));
  return MLF_OK;
}

// image_load_module_aug.m:128-148,160-162 (img_rectify_mod): query
// coordinates along one axis, step dx/M through `center`, from -dx/M to
// size + 0.9, started at a whole lenslet and cut to whole lenslets.  MATLAB's
// a:d:b is taken as a + k*d.
static std::vector<double> rectify_axis(double center, int64_t size, int M, double dx) {
  const double s = dx / M;
  auto colon = [](double a, double d, double b) {
    const int64_t n = (int64_t)std::floor((b - a) / d + 1e-10) + 1;
    std::vector<double> v;
    for (int64_t k = 0; k < n; ++k) v.push_back(a + d * (double)k);
    return v;
  };
  std::vector<double> first = colon(center, -s, -s);
  std::reverse(first.begin(), first.end());
  const std::vector<double> second = colon(center + s, s, (double)size + 0.9);
  std::vector<double> grid(first);
  grid.insert(grid.end(), second.begin(), second.end());
  const int64_t qc = (int64_t)first.size() - M / 2;  // find(grid == center) - floor(M/2)
  const int64_t q0 = qc - (int64_t)M * (int64_t)std::floor((double)qc / M);
  std::vector<double> q;
  if (q0 < (int64_t)grid.size()) q.assign(grid.begin() + q0, grid.end());
  q.resize(q.size() / M * M);
  return q;
}

static mlf_status rectify_check(int64_t H, int64_t W, double xc, double yc, int32_t M, double dx) {
  if (H < 2 || W < 2 || M < 1 || !(dx > 0) || !std::isfinite(xc) || !std::isfinite(yc) || xc < 0 ||
      yc < 0 || xc > (double)W || yc > (double)H)
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "rectification needs H, W >= 2, M >= 1, dx > 0 and a centre inside the image");
  return MLF_OK;
}

mlf_status mlf_rectify_size(int64_t H, int64_t W, double x_center, double y_center, int32_t M,
                            double dx, int64_t* out_h, int64_t* out_w) {
  clear_error();
  if (!out_h || !out_w) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  MLF_CHECK(rectify_check(H, W, x_center, y_center, M, dx));
  *out_h = (int64_t)rectify_axis(y_center, H, M, dx).size();
  *out_w = (int64_t)rectify_axis(x_center, W, M, dx).size();
  return MLF_OK;
}

mlf_status mlf_rectify_device(mlf_context ctx, const float* d_in, int64_t H, int64_t W, int64_t frames,
                              double x_center, double y_center, int32_t M, double dx, float* d_out) {
  clear_error();
  if (!ctx || !d_in || !d_out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  MLF_CHECK(rectify_check(H, W, x_center, y_center, M, dx));
  if (frames < 1 || H > 0x7fffffff || W > 0x7fffffff || frames > 0x7fffffff)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad frame count or size");
  const std::vector<double> yq = rectify_axis(y_center, H, M, dx), xq = rectify_axis(x_center, W, M, dx);
  if (yq.empty() || xq.empty() || xq.size() > 65535)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "rectified size %zu x %zu", yq.size(), xq.size());
  MLF_CHECK(ctx_activate(ctx));
  DevBuf<double> dq;
  MLF_CHECK(dq.alloc(yq.size() + xq.size()));
  MLF_HIP(hipMemcpyAsync(dq.ptr, yq.data(), sizeof(double) * yq.size(), hipMemcpyHostToDevice, ctx->stream));
  MLF_HIP(hipMemcpyAsync(dq.ptr + yq.size(), xq.data(), sizeof(double) * xq.size(), hipMemcpyHostToDevice,
                         ctx->stream));
  MLF_LAUNCH(launch_rectify(d_in, (int)H, (int)W, (int)frames, dq.ptr, (int)yq.size(), dq.ptr + yq.size(),
                            (int)xq.size(), d_out, ctx->stream));
  MLF_HIP(hipStreamSynchronize(ctx->stream));  // before the grid buffer is released
  return MLF_OK;
}

mlf_status mlf_rectify(const float* in, int64_t H, int64_t W, int64_t frames, double x_center,
                       double y_center, int32_t M, double dx, float* out, int32_t device) {
  clear_error();
  if (!in || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  int64_t hq = 0, wq = 0;
  MLF_CHECK(mlf_rectify_size(H, W, x_center, y_center, M, dx, &hq, &wq));
  if (frames < 1) return set_error(MLF_ERR_INVALID_ARGUMENT, "bad frame count");
  mlf_context ctx = nullptr;
  MLF_CHECK(mlf_context_create(device, &ctx));
  std::unique_ptr<mlf_context_s, mlf_status (*)(mlf_context)> guard(ctx, mlf_context_destroy);
  DevBuf<float> din, dout;
  MLF_CHECK(din.alloc((size_t)H * W * frames));
  MLF_CHECK(dout.alloc((size_t)hq * wq * frames));
  MLF_HIP(hipMemcpyAsync(din.ptr, in, din.bytes(), hipMemcpyHostToDevice, ctx->stream));
  MLF_CHECK(mlf_rectify_device(ctx, din.ptr, H, W, frames, x_center, y_center, M, dx, dout.ptr));
  MLF_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  MLF_HIP(hipStreamSynchronize(ctx->stream));
  return MLF_OK;
}

mlf_status mlf_carpet_patches(int64_t global_h, int64_t global_w, int32_t n_h, int32_t n_w,
                              int32_t M, int32_t min_size, int64_t* corners, int32_t* count) {
  clear_error();
  if (!corners || !count) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (global_h < 1 || global_w < 1 || n_h < 1 || n_w < 1 || M < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad image size, patch counts or M");
  // determine_patch_info_aug.m:106-112: floor(size / n), rounded up to a multiple of M
  const int64_t ph = (global_h / n_h + M - 1) / M * M, pw = (global_w / n_w + M - 1) / M * M;
  int32_t k = 0;
  for (int64_t i = 0; i < (int64_t)n_h * n_w; ++i) {
    const int64_t ci = i % n_h + 1, cj = i / n_h + 1;  // ind2sub, first index fastest (:115)
    const int64_t top = (ci - 1) * ph + 1, left = (cj - 1) * pw + 1;
    const int64_t bottom = std::min(ci * ph, global_h), right = std::min(cj * pw, global_w);
    int64_t h = bottom - top + 1, w = right - left + 1;
    h -= h % M;  // :126-127
    w -= w % M;
    if (h < min_size || w < min_size) continue;  // :129-132, discarded
    int64_t* c = corners + 4 * (size_t)k++;
    c[0] = top;
    c[1] = left;
    c[2] = top + h - 1;
    c[3] = left + w - 1;
  }
  *count = k;
  return MLF_OK;
}

// 'roi' tiling (determine_patch_info_aug.m:144-218): each ROI [top, left,
// bottom, right] is tiled with patches of the patch size cut down to a
// multiple of M (:35-41); an ROI smaller than a patch along an axis gives one
// patch of the ROI's size there (:170-183).
static void roi_tiles(const int64_t* roi, int32_t n_roi, int32_t roi_patch_size, int M, int min_size,
                      std::vector<std::array<int64_t, 4>>* out) {
  const int64_t ps = roi_patch_size - roi_patch_size % M;
  for (int32_t q = 0; q < n_roi; ++q) {
    const int64_t* c = roi + 4 * (size_t)q;
    const int64_t sh = c[2] - c[0] + 1, sw = c[3] - c[1] + 1;
    int64_t nh = (sh + ps - 1) / ps, nw = (sw + ps - 1) / ps, ph = ps, pw = ps;
    if (sh < ps) ph = sh, nh = 1;
    if (sw < ps) pw = sw, nw = 1;
    for (int64_t j = 0; j < nh * nw; ++j) {
      const int64_t ci = j % nh + 1, cj = j / nh + 1;
      const int64_t top = (ci - 1) * ph + c[0], left = (cj - 1) * pw + c[1];
      const int64_t bottom = std::min(ci * ph, sh) + c[0] - 1, right = std::min(cj * pw, sw) + c[1] - 1;
      int64_t h = bottom - top + 1, w = right - left + 1;
      h -= h % M;
      w -= w % M;
      if (h < min_size || w < min_size) continue;  // :201-204, discarded
      out->push_back({top, left, top + h - 1, left + w - 1});
    }
  }
}

// a:d:b as a + k*d (also rectify_axis's convention)
static std::vector<double> colon_seq(double a, double d, double b) {
  const int64_t n = (int64_t)std::floor((b - a) / d + 1e-10) + 1;
  std::vector<double> v;
  for (int64_t k = 0; k < n; ++k) v.push_back(a + d * (double)k);
  return v;
}

static int64_t matlab_round_i(double x) { return (int64_t)(x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5)); }

mlf_status mlf_determine_patch_info(int64_t raw_h, int64_t raw_w, double x_center, double y_center,
                                   double dx, int32_t M, const int32_t* crop_border, int32_t mode,
                                   const int32_t* n_patches, const int64_t* roi_list, int32_t n_roi,
                                   int32_t roi_patch_size, mlf_patch_info* out, int32_t capacity,
                                   int32_t* count) {
  clear_error();
  if (!crop_border || !count) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (raw_h < 2 || raw_w < 2 || M < 1 || !(dx > 0.0) || !std::isfinite(dx))
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad frame size, M or dx");
  if (!(x_center >= 1.0 && x_center <= (double)raw_w && y_center >= 1.0 && y_center <= (double)raw_h))
    return set_error(MLF_ERR_INVALID_ARGUMENT, "frame centre outside the frame");
  const double s = dx / M;
  // reference grid of the whole frame (:51-52): step dx/M through the centre, 1 .. size
  std::vector<double> xf = colon_seq(x_center, -s, 1.0), yf = colon_seq(y_center, -s, 1.0);
  const int64_t pre_x = (int64_t)xf.size(), pre_y = (int64_t)yf.size();  // find(Xresample == xCenter)
  std::reverse(xf.begin(), xf.end());
  std::reverse(yf.begin(), yf.end());
  std::vector<double> Xr(xf), Yr(yf);
  const std::vector<double> xs = colon_seq(x_center + s, s, (double)raw_w), ys = colon_seq(y_center + s, s, (double)raw_h);
  Xr.insert(Xr.end(), xs.begin(), xs.end());
  Yr.insert(Yr.end(), ys.begin(), ys.end());
  // started at a whole lenslet (:60-69), cut to whole lenslets (:82-84)
  const int64_t Md = M / 2;
  const int64_t xq = (pre_x - Md) - (int64_t)M * (int64_t)std::floor((double)(pre_x - Md) / M);
  const int64_t yq = (pre_y - Md) - (int64_t)M * (int64_t)std::floor((double)(pre_y - Md) / M);
  const int64_t w_res = (int64_t)M * (((int64_t)Xr.size() - xq) / M);
  const int64_t h_res = (int64_t)M * (((int64_t)Yr.size() - yq) / M);
  // border crop in whole microlenses (:87-103)
  const int64_t cl = crop_border[0], cr = crop_border[1], cu = crop_border[2], cd = crop_border[3];
  if (cl < 0 || cr < 0 || cu < 0 || cd < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "negative crop");
  if (cl + cr >= w_res / M) return set_error(MLF_ERR_INVALID_ARGUMENT, "X-cut range is larger than the x-size of image");
  if (cu + cd >= h_res / M) return set_error(MLF_ERR_INVALID_ARGUMENT, "Y-cut range is larger than the y-size of image");
  const int64_t x_first = 1 + cl, y_first = 1 + cu;
  const int64_t gw = (w_res / M - cr - x_first + 1) * M, gh = (h_res / M - cd - y_first + 1) * M;
  std::vector<std::array<int64_t, 4>> tiles;
  if (mode == MLF_PATCH_CARPET) {
    if (!n_patches) return set_error(MLF_ERR_INVALID_ARGUMENT, "n_patches is NULL");
    std::vector<int64_t> c(4 * (size_t)std::max(1, n_patches[0]) * std::max(1, n_patches[1]));
    int32_t n = 0;
    MLF_CHECK(mlf_carpet_patches(gh, gw, n_patches[0], n_patches[1], M, 50, c.data(), &n));
    for (int32_t q = 0; q < n; ++q) tiles.push_back({c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]});
  } else if (mode == MLF_PATCH_ROI) {
    if (!roi_list || n_roi < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "ROI list missing");
    if (roi_patch_size < M) return set_error(MLF_ERR_INVALID_ARGUMENT, "ROI patch size %d < M", (int)roi_patch_size);
    for (int32_t q = 0; q < n_roi; ++q) {
      const int64_t* c = roi_list + 4 * (size_t)q;
      if (c[0] < 1 || c[1] < 1 || c[2] < c[0] || c[3] < c[1])
        return set_error(MLF_ERR_INVALID_ARGUMENT, "ROI %d is not [top, left, bottom, right]", (int)q);
    }
    roi_tiles(roi_list, n_roi, roi_patch_size, M, 50, &tiles);
  } else {
    return set_error(MLF_ERR_INVALID_ARGUMENT, "patch_mode should be either 'carpet' or 'roi'");
  }
  *count = (int32_t)tiles.size();
  if (!out) return MLF_OK;  // count query
  if ((int32_t)tiles.size() > capacity)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "%zu patches, capacity %d", tiles.size(), (int)capacity);
  // raw cut and microlens centre per patch (:221-275)
  const int64_t y_top = (y_first - 1) * M + 1, x_left = (x_first - 1) * M + 1;
  const int64_t margin = (M + 1) / 2 - 2;  // ceil(M/2) - 2
  const int64_t half = (M + 1) / 2;
  for (size_t q = 0; q < tiles.size(); ++q) {
    const auto& t = tiles[q];
    const int64_t pt = t[0] + y_top - 1 + yq - 1 - margin, pl = t[1] + x_left - 1 + xq - 1 - margin;
    const int64_t pb = t[2] + y_top - 1 + yq - 1 + margin - 1, pr = t[3] + x_left - 1 + xq - 1 + margin - 1;
    if (pt < 1 || pl < 1 || pb > (int64_t)Yr.size() || pr > (int64_t)Xr.size())
      return set_error(MLF_ERR_INVALID_ARGUMENT,
                       "patch %zu: its raw cut leaves the frame's resample grid (the reference stops "
                       "there too, determine_patch_info_aug.m:245-248); crop more border microlenses",
                       q + 1);
    mlf_patch_info& o = out[q];
    o.location[0] = t[0];
    o.location[1] = t[1];
    o.location[2] = t[2];
    o.location[3] = t[3];
    o.size[0] = t[2] - t[0] + 1;
    o.size[1] = t[3] - t[1] + 1;
    o.raw_cut[0] = std::max<int64_t>((int64_t)std::floor(Yr[pt - 1]), 1);
    o.raw_cut[1] = std::min<int64_t>((int64_t)std::ceil(Yr[pb - 1]), raw_h);
    o.raw_cut[2] = std::max<int64_t>((int64_t)std::floor(Xr[pl - 1]), 1);
    o.raw_cut[3] = std::min<int64_t>((int64_t)std::ceil(Xr[pr - 1]), raw_w);
    const int64_t cy = matlab_round_i((double)(pt + pb) / 2.0), cx = matlab_round_i((double)(pl + pr) / 2.0);
    const int64_t my = ((pre_y - cy) % M + M) % M, mx = ((pre_x - cx) % M + M) % M;
    const int64_t iy = my >= half ? cy - (M - my) : cy + my;
    const int64_t ix = mx >= half ? cx - (M - mx) : cx + mx;
    if (iy < 1 || ix < 1 || iy > (int64_t)Yr.size() || ix > (int64_t)Xr.size())
      return set_error(MLF_ERR_INVALID_ARGUMENT, "patch %zu: centre outside the resample grid", q + 1);
    o.raw_cut_yCenter = Yr[iy - 1] - Yr[pt - 1];
    o.raw_cut_xCenter = Xr[ix - 1] - Xr[pl - 1];
  }
  return MLF_OK;
}

mlf_status mlf_crop_movie_device(mlf_context ctx, const float* d_movie, int64_t H, int64_t W,
                                 int64_t frames, int64_t top, int64_t left, int64_t h, int64_t w,
                                 float* d_out) {
  clear_error();
  if (!ctx || !d_movie || !d_out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (H < 1 || W < 1 || frames < 1 || frames > 0x7fffffff || h < 1 || w < 1 || h > 0x7fffffff ||
      w > 65535 || top < 1 || left < 1 || top + h - 1 > H || left + w - 1 > W)
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "patch rows %lld..%lld, columns %lld..%lld outside the %lld x %lld movie",
                     (long long)top, (long long)(top + h - 1), (long long)left,
                     (long long)(left + w - 1), (long long)H, (long long)W);
  MLF_CHECK(ctx_activate(ctx));
  MLF_LAUNCH(launch_crop_movie(d_movie, H, W, top - 1, left - 1, (int)h, (int)w, (int)frames, d_out,
                               ctx->stream));
  return MLF_OK;
}

static mlf_status summary_device(mlf_context ctx, const float* d_movie, int64_t height,
                                 int64_t width, int64_t frames, int32_t window_1, int32_t window_2,
                                 int32_t poly_index, int32_t bg_iter, int global, double* d_summary,
                                 double* d_bg_spatial, double* d_bg_temporal);

mlf_status mlf_summary_image_device(mlf_context ctx, const float* d_movie, int64_t height,
                                   int64_t width, int64_t frames, int32_t window_1,
                                   int32_t window_2, int32_t poly_index, int32_t bg_iter,
                                   double* d_summary, double* d_bg_spatial, double* d_bg_temporal) {
  clear_error();
  return summary_device(ctx, d_movie, height, width, frames, window_1, window_2, poly_index, bg_iter,
                        0, d_summary, d_bg_spatial, d_bg_temporal);
}

mlf_status mlf_summary_image_global_device(mlf_context ctx, const float* d_movie, int64_t height,
                                          int64_t width, int64_t frames, int32_t bg_iter,
                                          double* d_summary, double* d_bg_spatial,
                                          double* d_bg_temporal) {
  clear_error();
  return summary_device(ctx, d_movie, height, width, frames, 1, 1, 0, bg_iter, 1, d_summary,
                        d_bg_spatial, d_bg_temporal);
}

static mlf_status summary_device(mlf_context ctx, const float* d_movie, int64_t height,
                                 int64_t width, int64_t frames, int32_t window_1, int32_t window_2,
                                 int32_t poly_index, int32_t bg_iter, int global, double* d_summary,
                                 double* d_bg_spatial, double* d_bg_temporal) {
  if (!ctx || !d_movie || !d_summary) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (height < 1 || width < 1 || frames < 1 || frames > 0x7fffffff)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad movie dims %lld x %lld x %lld", (long long)height,
                     (long long)width, (long long)frames);
  // mesolf_set_params.m:120-132: windows and the exponent are non-negative
  // integers; a window of 0 selects nothing, and bg_iter = 0 leaves the
  // temporal background undefined in rank_1_factorization.m
  if (window_1 < 1 || window_2 < 1 || poly_index < 0 || bg_iter < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "window_1, window_2 and bg_iter must be >= 1 and poly_index >= 0");
  MLF_CHECK(ctx_activate(ctx));
  const size_t npix = (size_t)height * width;
  const int T = (int)frames;
  DevBuf<double> scratch, s_own;
  DevBuf<float> dbuf;
  MLF_CHECK(scratch.alloc(summary_scratch_doubles(npix, T)));
  if (!global) MLF_CHECK(dbuf.alloc(npix * (size_t)T));
  double* s = d_bg_spatial;
  if (!s) {
    MLF_CHECK(s_own.alloc(npix));
    s = s_own.ptr;
  }
  MLF_LAUNCH(launch_summary(d_movie, npix, T, window_1, window_2, poly_index, bg_iter, global,
                            scratch.ptr, dbuf.ptr, d_summary, s, ctx->stream));
  if (d_bg_temporal)
    MLF_HIP(hipMemcpyAsync(d_bg_temporal, summary_temporal(scratch.ptr), sizeof(double) * T,
                           hipMemcpyDeviceToDevice, ctx->stream));
  MLF_HIP(hipStreamSynchronize(ctx->stream));  // before the scratch is released
  return MLF_OK;
}

static mlf_status summary_host(const float* movie, int64_t height, int64_t width, int64_t frames,
                               int32_t window_1, int32_t window_2, int32_t poly_index,
                               int32_t bg_iter, int global, double* summary, double* bg_spatial,
                               double* bg_temporal, int32_t device);

mlf_status mlf_summary_image(const float* movie, int64_t height, int64_t width, int64_t frames,
                             int32_t window_1, int32_t window_2, int32_t poly_index,
                             int32_t bg_iter, double* summary, double* bg_spatial,
                             double* bg_temporal, int32_t device) {
  clear_error();
  return summary_host(movie, height, width, frames, window_1, window_2, poly_index, bg_iter, 0,
                      summary, bg_spatial, bg_temporal, device);
}

mlf_status mlf_summary_image_global(const float* movie, int64_t height, int64_t width,
                                    int64_t frames, int32_t bg_iter, double* summary,
                                    double* bg_spatial, double* bg_temporal, int32_t device) {
  clear_error();
  return summary_host(movie, height, width, frames, 1, 1, 0, bg_iter, 1, summary, bg_spatial,
                      bg_temporal, device);
}

static mlf_status summary_host(const float* movie, int64_t height, int64_t width, int64_t frames,
                               int32_t window_1, int32_t window_2, int32_t poly_index,
                               int32_t bg_iter, int global, double* summary, double* bg_spatial,
                               double* bg_temporal, int32_t device) {
  if (!movie || !summary) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (height < 1 || width < 1 || frames < 1 || frames > 0x7fffffff)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad movie dims");
  mlf_context ctx = nullptr;
  MLF_CHECK(mlf_context_create(device, &ctx));
  std::unique_ptr<mlf_context_s, mlf_status (*)(mlf_context)> guard(ctx, mlf_context_destroy);
  const size_t npix = (size_t)height * width;
  DevBuf<float> dm;
  DevBuf<double> dimg, ds, dt;
  MLF_CHECK(dm.alloc(npix * (size_t)frames));
  MLF_CHECK(dimg.alloc(npix));
  MLF_CHECK(ds.alloc(npix));
  MLF_CHECK(dt.alloc((size_t)frames));
  MLF_HIP(hipMemcpyAsync(dm.ptr, movie, dm.bytes(), hipMemcpyHostToDevice, ctx->stream));
  MLF_CHECK(summary_device(ctx, dm.ptr, height, width, frames, window_1, window_2, poly_index,
                           bg_iter, global, dimg.ptr, ds.ptr, dt.ptr));
  MLF_HIP(hipMemcpyAsync(summary, dimg.ptr, dimg.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  if (bg_spatial)
    MLF_HIP(hipMemcpyAsync(bg_spatial, ds.ptr, ds.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  if (bg_temporal)
    MLF_HIP(hipMemcpyAsync(bg_temporal, dt.ptr, dt.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  MLF_HIP(hipStreamSynchronize(ctx->stream));
  return MLF_OK;
}

static mlf_status segments_check(int64_t width, int64_t height, int64_t nslices, const int32_t* plane,
                                 const int64_t* seg_ptr, int64_t nseg, int64_t kw, int64_t nphase,
                                 int64_t Z) {
  if (!plane || !seg_ptr) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width < 1 || height < 1 || nslices < 0 || nseg < 1 || Z < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad dimensions");
  if (nphase != kN) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernels must be [kw,kw,15,15,Z]");
  if (kw < 1 || kw > 2048) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernel width out of [1,2048]");
  if (nseg > 65535) return set_error(MLF_ERR_INVALID_ARGUMENT, "too many segments (> 65535)");
  if (nslices > INT32_MAX / 2) return set_error(MLF_ERR_INVALID_ARGUMENT, "too many slices");
  if (seg_ptr[0] != 0 || seg_ptr[nseg] != nslices)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "seg_ptr must start at 0 and end at nslices");
  for (int64_t s = 0; s < nseg; ++s)
    if (seg_ptr[s + 1] < seg_ptr[s]) return set_error(MLF_ERR_INVALID_ARGUMENT, "seg_ptr decreases");
  for (int64_t i = 0; i < nslices; ++i)
    if (plane[i] < 0 || plane[i] >= Z)
      return set_error(MLF_ERR_INVALID_ARGUMENT, "slice %lld: plane %d outside [0,%lld)", (long long)i,
                       (int)plane[i], (long long)Z);
  return MLF_OK;
}

// The seed-library batch on device buffers (host plane / seg_ptr tables).
// Every segment's contributions are added in ascending plane order, slices of
// one plane in slice order: the order of the reference's loop over planes
// (gen_seed_library_local_with_ring_bg_same_size.m:263-318).  Odd kw: per
// plane, the kernel's GEMM operand is built once and the plane's slices run
// through the MFMA operator in rounds whose output segments are distinct,
// each adding its projection into its segment.  Even kw: the scalar kernel
// walks each segment's slices in the same order.
static mlf_status segments_device(mlf_context ctx, const float* d_in, int64_t W, int64_t H,
                                  int64_t nslices, const int32_t* plane, const int64_t* seg_ptr,
                                  int64_t nseg, const float* d_K, int kw, int64_t Z, float* d_out) {
  hipStream_t st = ctx->stream;
  const size_t img = (size_t)W * H;
  std::vector<int32_t> seg_of(nslices);
  for (int64_t s = 0; s < nseg; ++s)
    for (int64_t i = seg_ptr[s]; i < seg_ptr[s + 1]; ++i) seg_of[i] = (int32_t)s;
  const size_t kplane = (size_t)kw * kw * kP;
  if (kw % 2 == 1 && kw <= kMexMfmaMaxKw) {
    // launches: (plane, round) -> lists of (slice, segment)
    std::vector<std::vector<int32_t>> by_plane(Z);
    for (int64_t i = 0; i < nslices; ++i) by_plane[plane[i]].push_back((int32_t)i);
    std::vector<int32_t> src, dst;
    struct Launch { int64_t z, off, n; };
    std::vector<Launch> launches;
    std::vector<int32_t> round_of_seg(nseg, 0);
    for (int64_t z = 0; z < Z; ++z) {
      if (by_plane[z].empty()) continue;
      std::vector<std::vector<int32_t>> rounds;
      for (int32_t i : by_plane[z]) {
        const int r = round_of_seg[seg_of[i]]++;
        if ((int)rounds.size() <= r) rounds.resize(r + 1);
        rounds[r].push_back(i);
      }
      for (int32_t i : by_plane[z]) round_of_seg[seg_of[i]] = 0;
      for (const auto& rd : rounds) {
        launches.push_back(Launch{z, (int64_t)src.size(), (int64_t)rd.size()});
        for (int32_t i : rd) {
          src.push_back(i);
          dst.push_back(seg_of[i]);
        }
      }
    }
    DevBuf<int32_t> d_src, d_dst;
    MLF_CHECK(upload_vec(d_src, src));
    MLF_CHECK(upload_vec(d_dst, dst));
    MLF_LAUNCH(launch_fill(d_out, img * nseg, 0.f, st));
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t need = mex_kq_floats(kw);
    if (ctx->mex_kq.count < need) {
      MLF_HIP(hipStreamSynchronize(st));  // earlier calls may still read it
      MLF_CHECK(ctx->mex_kq.alloc(need));
    }
    int64_t zprev = -1;
    for (const Launch& l : launches) {
      if (l.z != zprev) MLF_LAUNCH(launch_mex_kq(d_K + (size_t)l.z * kplane, kw, ctx->mex_kq.ptr, st));
      zprev = l.z;
      MLF_LAUNCH(launch_mex_mfma_list(d_in, W, H, l.n, ctx->mex_kq.ptr, kw, d_src.ptr + l.off,
                                      d_dst.ptr + l.off, d_out, st));
    }
    // the index lists are freed on return: wait for the launches that read them
    MLF_HIP(hipStreamSynchronize(st));
    return MLF_OK;
  }
  std::vector<int32_t> order(nslices);
  for (int64_t s = 0; s < nseg; ++s) {
    for (int64_t i = seg_ptr[s]; i < seg_ptr[s + 1]; ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin() + seg_ptr[s], order.begin() + seg_ptr[s + 1],
                     [&](int32_t a, int32_t b) { return plane[a] < plane[b]; });
  }
  DevBuf<int32_t> dplane, dorder;
  DevBuf<int64_t> dptr;
  MLF_CHECK(upload_vec(dplane, std::vector<int32_t>(plane, plane + nslices)));
  MLF_CHECK(upload_vec(dorder, order));
  MLF_CHECK(upload_vec(dptr, std::vector<int64_t>(seg_ptr, seg_ptr + nseg + 1)));
  MLF_LAUNCH(launch_mex_conv_segments(d_in, W, H, dplane.ptr, dptr.ptr, dorder.ptr, nseg, d_K, kw,
                                      d_out, st));
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

mlf_status mlf_lfm_convolution_segments(const float* slices, int64_t width, int64_t height,
                                        int64_t nslices, const int32_t* plane,
                                        const int64_t* seg_ptr, int64_t nseg,
                                        const float* kernels, int64_t kw, int64_t nphase,
                                        int64_t Z, float* out, int32_t device) {
  clear_error();
  if (!kernels || !out || (nslices > 0 && !slices))
    return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  MLF_CHECK(segments_check(width, height, nslices, plane, seg_ptr, nseg, kw, nphase, Z));
  mlf_context ctx = nullptr;
  MLF_CHECK(mlf_context_create(device, &ctx));
  std::unique_ptr<mlf_context_s, mlf_status (*)(mlf_context)> guard(ctx, mlf_context_destroy);
  hipStream_t st = ctx->stream;
  const size_t img = (size_t)width * height;
  DevBuf<float> din, dk, dout;
  MLF_CHECK(din.alloc(std::max<size_t>(1, img * nslices)));
  MLF_CHECK(dk.alloc((size_t)kw * kw * kP * Z));
  MLF_CHECK(dout.alloc(img * nseg));
  if (nslices > 0)
    MLF_HIP(hipMemcpyAsync(din.ptr, slices, img * nslices * sizeof(float), hipMemcpyHostToDevice, st));
  MLF_HIP(hipMemcpyAsync(dk.ptr, kernels, dk.bytes(), hipMemcpyHostToDevice, st));
  MLF_CHECK(segments_device(ctx, din.ptr, width, height, nslices, plane, seg_ptr, nseg, dk.ptr, (int)kw,
                            Z, dout.ptr));
  MLF_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, st));
  MLF_HIP(hipStreamSynchronize(st));
  return MLF_OK;
}

mlf_status mlf_lfm_convolution_segments_device(mlf_context ctx, const float* d_slices, int64_t width,
                                               int64_t height, int64_t nslices, const int32_t* plane,
                                               const int64_t* seg_ptr, int64_t nseg,
                                               const float* d_kernels, int64_t kw, int64_t nphase,
                                               int64_t Z, float* d_out) {
  clear_error();
  if (!ctx || !d_kernels || !d_out || (nslices > 0 && !d_slices))
    return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  MLF_CHECK(segments_check(width, height, nslices, plane, seg_ptr, nseg, kw, nphase, Z));
  MLF_CHECK(ctx_activate(ctx));
  return segments_device(ctx, d_slices, width, height, nslices, plane, seg_ptr, nseg, d_kernels, (int)kw,
                         Z, d_out);
}

}  // extern "C"
