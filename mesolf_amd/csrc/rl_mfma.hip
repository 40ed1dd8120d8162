// MFMA implicit-GEMM kernels of the phase-space RL view step (gfx950).
//
// Forward projection of one view (recon:109, 243-249), per plane z:
//     Q[L, e] = sum_phase X[z][L][phase] * W[z][e][phase]        (GEMM, K = phase)
//     y[i]   += Q[i - e, e]                                      (diagonal scatter)
//   A operand = the volume block (registers, loaded once from HBM), B operand =
//   the PSF tap rows (LDS, double-buffered chunks of 64 taps), the scatter
//   goes into an LDS window that is stored as a per-(plane, block) partial;
//   k_ratio_gather sums the partials in a fixed order (deterministic).
//
// Backprojection + multiplicative update (recon:121-137), per plane z:
//     B[a, phase] = sum_e R[a - e] * Wt[z][e][phase]             (GEMM, K = taps)
//   A operand = im2col of the ratio image (LDS window), B operand = the
//   transposed PSF table (global, taps contiguous).  The backprojection of
//   ones (the normaliser) is read from a precomputed FP64-summed table.
//
// 16x16 MFMA operand layout used throughout (lane l, r = l % 16, g = l / 16):
//   A: row r, k = g*KPL + j;  B: column r, k = g*KPL + j;  C: rows 4g + j, column r.
#include "internal.h"
#include "rl_mfma.h"

namespace mlf {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Mma;

template <>
struct Mma<float> {
  static constexpr int KSTEP = 4, KPL = 1;
  static constexpr int PHI = kPhiPadF32;   // K extent of the forward GEMM
  static constexpr int LDB = kPhiPadF32;   // LDS row stride of forward B chunks
  typedef float frag;
  __device__ static __forceinline__ floatx4 mma(frag a, frag b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ frag load(const float* p) { return *p; }
};

template <>
struct Mma<__bf16> {
  static constexpr int KSTEP = 32, KPL = 8;
  static constexpr int PHI = kPhiPadBF16;
  static constexpr int LDB = kPhiPadBF16;  // XOR-swizzled 16-byte columns
  typedef bf16x8 frag;
  __device__ static __forceinline__ floatx4 mma(frag a, frag b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ frag load(const __bf16* p) { return *(const bf16x8*)p; }
};

__device__ __forceinline__ int ceil_div(int a, int b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------ forward
template <typename T, int MT>
__global__ __launch_bounds__(256) void k_fwd_mfma(const float* __restrict__ X,
                                                  const T* __restrict__ W, FwdGeo g,
                                                  float* __restrict__ partial) {
  using M = Mma<T>;
  constexpr int KSTEP = M::KSTEP, KPL = M::KPL, PHI = M::PHI, LDB = M::LDB;
  constexpr int NKS = PHI / KSTEP;
  constexpr int BM2 = 4 * MT;
  constexpr int CH = kFwdChunk;  // taps per chunk (4 N-tiles of 16)
  extern __shared__ __align__(16) char smem[];
  T* Bs = reinterpret_cast<T*>(smem);                                  // [2][CH][LDB]
  float* ywin = reinterpret_cast<float*>(smem + sizeof(T) * 2 * CH * LDB);

  const int* s = g.sched + 3 * blockIdx.x;
  const int z = s[0], b1 = s[1], b2 = s[2];
  const int r = g.r[z];
  const int K = 2 * r + 1, ntaps = K * K;
  const int nch = g.nchunk[z];
  const int W1 = 16 + 2 * r, W2 = BM2 + 2 * r;
  const int WW = MT + 2 * r;  // rows of one wave's window
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lg = lane >> 4;
  // one window per wave: atomics of a single wave land in program/lane order,
  // and the four windows are summed in a fixed order -> deterministic partials
  float* yw = ywin + w * W1 * WW;

  for (int o = tid; o < 4 * W1 * WW; o += 256) ywin[o] = 0.f;

  // A operand: this lane's lenslet rows of the volume block, all phases.
  typename M::frag a[MT][NKS];
  {
    const int a1 = 16 * b1 + lr;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int a2 = BM2 * b2 + MT * w + mt;
      const bool ok = a1 < g.n1 && a2 < g.n2;
      const float* xp = X + (((size_t)z * g.n2 + (ok ? a2 : 0)) * g.n1 + (ok ? a1 : 0)) * kP;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if constexpr (KPL == 1) {
          const int ph = ks * KSTEP + lg;
          a[mt][ks] = (ok && ph < kP) ? xp[ph] : 0.f;
        } else {
          // 8 consecutive phases; a lenslet row is 900 B (4 B aligned only)
          const int ph0 = ks * KSTEP + lg * KPL;
          float v[KPL];
          if (ok && ph0 + KPL <= kP) {
#pragma unroll
            for (int j = 0; j < KPL; ++j) v[j] = xp[ph0 + j];
          } else {
#pragma unroll
            for (int j = 0; j < KPL; ++j) v[j] = (ok && ph0 + j < kP) ? xp[ph0 + j] : 0.f;
          }
#pragma unroll
          for (int j = 0; j < KPL; ++j) a[mt][ks][j] = (__bf16)v[j];
        }
      }
    }
  }

  const T* Wz = W + g.woff[z];
  // Chunk c (CH rows of PHI elements) is copied global -> LDS with
  // global_load_lds_dwordx4 (no staging registers), into the buffer the
  // waves are not reading, while the MFMAs of chunk c-1 run.  Each wave pass
  // fills 64 consecutive 16-byte LDS slots; for BF16 the slot of logical
  // vector (row, col) is row*RV + (col ^ (row & 31)) so that the sixteen rows
  // read by one ds_read_b128 land in distinct banks.
  constexpr int VEC = 16 / sizeof(T);            // elements per 16 B
  constexpr int RV = PHI / VEC;                  // vectors per row
  constexpr bool SWZ = (RV == 32);
  constexpr int NSLOT = CH * RV;
  auto issue_chunk = [&](int c, int buf) {
    const T* src = Wz + (size_t)c * CH * PHI;
    char* dst = reinterpret_cast<char*>(Bs) + (size_t)buf * CH * LDB * sizeof(T);
    for (int base = w * 64; base < NSLOT; base += 256) {
      const int slot = base + lane;
      if (slot < NSLOT) {
        const int row = slot / RV, pos = slot - row * RV;
        const int col = SWZ ? (pos ^ (row & 31)) : pos;
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(src + (size_t)row * PHI + col * VEC),
            (__attribute__((address_space(3))) void*)(dst + (size_t)base * 16), 16, 0, 0);
      }
    }
  };

  issue_chunk(0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  for (int c = 0; c < nch; ++c) {
    __syncthreads();  // chunk c landed; everyone done reading the other buffer
    if (c + 1 < nch) issue_chunk(c + 1, (c + 1) & 1);
    const T* Bc = Bs + (size_t)(c & 1) * CH * LDB;
    floatx4 acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      typename M::frag bf[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const int row = nt * 16 + lr;
        if constexpr (SWZ) {
          const int col = (ks * KSTEP + lg * KPL) / VEC;
          bf[nt] = M::load(Bc + row * LDB + ((col ^ (row & 31)) * VEC));
        } else {
          bf[nt] = M::load(Bc + row * LDB + ks * KSTEP + lg * KPL);
        }
      }
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = M::mma(a[mt][ks], bf[nt], acc[mt][nt]);
    }
    // diagonal scatter into the LDS window: y[a + d] += Q[a, e(d)]
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int e = c * CH + nt * 16 + lr;
      if (e >= ntaps) continue;
      const int e1 = e / K, e2 = e - e1 * K;  // packed rows are e2-fastest
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        float* row = yw + (mt + e2) * W1 + e1;
#pragma unroll
        for (int j = 0; j < 4; ++j) atomicAdd(row + 4 * lg + j, acc[mt][nt][j]);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of chunk c+1 done
  }
  __syncthreads();
  const int nb1 = ceil_div(g.n1, 16);
  float* out = partial + g.poff[z] + (size_t)(b2 * nb1 + b1) * W1 * W2;
  for (int o = tid; o < W1 * W2; o += 256) {
    const int y = o / W1, x = o - y * W1;
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int yy = y - MT * q;
      if (yy >= 0 && yy < WW) acc += ywin[(q * WW + yy) * W1 + x];
    }
    out[o] = acc;
  }
}

// y[i] = sum over planes and covering blocks of the window partials (fixed
// order, FP64), then R = LF(u::15, v::15) ./ y with non-finite -> 0.
__global__ void k_ratio_gather(const float* __restrict__ partial, FwdGeo g, int Z, int BM2,
                               const float* __restrict__ LF, int u, int v, float* __restrict__ R,
                               float* __restrict__ y_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n1 * g.n2) return;
  const int i1 = i % g.n1, i2 = i / g.n1;
  const int nb1 = ceil_div(g.n1, 16), nb2 = ceil_div(g.n2, BM2);
  double y = 0.0;
  for (int z = 0; z < Z; ++z) {
    const int r = g.r[z];
    const int W1 = 16 + 2 * r, W2 = BM2 + 2 * r;
    const int b1h = min(nb1 - 1, (i1 + r) / 16);
    const int b1l = (i1 - r - 15 <= 0) ? 0 : (i1 - r - 15 + 15) / 16;
    const int b2h = min(nb2 - 1, (i2 + r) / BM2);
    const int b2l = (i2 - r - BM2 + 1 <= 0) ? 0 : (i2 - r - BM2 + 1 + BM2 - 1) / BM2;
    const float* pz = partial + g.poff[z];
    for (int b2 = b2l; b2 <= b2h; ++b2)
      for (int b1 = b1l; b1 <= b1h; ++b1)
        y += (double)pz[(size_t)(b2 * nb1 + b1) * W1 * W2 + (i2 - BM2 * b2 + r) * W1 +
                        (i1 - 16 * b1 + r)];
  }
  const float yf = (float)y;
  if (LF) {
    const float lf = LF[(size_t)(u + kN * i1) + (size_t)(kN * g.n1) * (v + kN * i2)];
    float q = lf / yf;
    if (!isfinite(q)) q = 0.f;
    R[i] = q;
  }
  if (y_out) y_out[i] = yf;
}

// --------------------------------------------------------------- backward
__device__ __forceinline__ int norm_class(int a, int n, int r) {
  if (n < 2 * r + 1) return a;  // small grid: one class per lenslet
  return a < r ? a : (a > n - 1 - r ? 2 * r - (n - 1 - a) : r);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_mfma(float* __restrict__ X, const T* __restrict__ Wb,
                                                  const float* __restrict__ Tn, BwdGeo g,
                                                  const float* __restrict__ R, float wgt) {
  using M = Mma<T>;
  constexpr int KSTEP = M::KSTEP, KPL = M::KPL;
  constexpr int MT = 4, NT = kBwdNTiles;
  extern __shared__ __align__(16) char smem[];
  const int* s = g.sched + 4 * blockIdx.x;
  const int z = s[0], b1 = s[1], b2 = s[2], grp = s[3];
  const int r = g.r[z];
  const int K = 2 * r + 1, ntaps = K * K;
  const int epad = g.epad[z];
  const int Wd = 16 + 2 * r;
  T* Rw = reinterpret_cast<T*>(smem);                               // [Wd][Wd]
  int* toff = reinterpret_cast<int*>(smem + sizeof(T) * ((Wd * Wd + 7) & ~7));  // [epad]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lg = lane >> 4;

  // ratio window (zero outside the grid) and tap offsets
  for (int o = tid; o < Wd * Wd; o += 256) {
    const int x = o % Wd, y = o / Wd;
    const int a1 = 16 * b1 - r + x, a2 = 16 * b2 - r + y;
    const float val = (a1 >= 0 && a1 < g.n1 && a2 >= 0 && a2 < g.n2) ? R[a1 + g.n1 * a2] : 0.f;
    Rw[o] = (T)val;
  }
  for (int e = tid; e < epad; e += 256) {
    int off = 0;  // padded taps have zero weights; any in-window read is fine
    if (e < ntaps) {
      const int e2 = e / K, e1 = e - e2 * K;
      off = -(e1 - r) - (e2 - r) * Wd;
    }
    toff[e] = off;
  }
  __syncthreads();

  int base[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) base[mt] = (lr + r) + (4 * w + mt + r) * Wd;

  floatx4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};

  const T* Bz = Wb + g.boff[z] + (size_t)(grp * 16 * NT) * epad;
  const int nks = epad / KSTEP;
  typename M::frag bn[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) bn[nt] = M::load(Bz + (size_t)(nt * 16 + lr) * epad + lg * KPL);
  for (int ks = 0; ks < nks; ++ks) {
    typename M::frag bf[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bf[nt] = bn[nt];
    if (ks + 1 < nks) {
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        bn[nt] = M::load(Bz + (size_t)(nt * 16 + lr) * epad + (ks + 1) * KSTEP + lg * KPL);
    }
    int o[KPL];
#pragma unroll
    for (int j = 0; j < KPL; ++j) o[j] = toff[ks * KSTEP + lg * KPL + j];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      typename M::frag af;
      if constexpr (KPL == 1) {
        af = Rw[base[mt] + o[0]];
      } else {
#pragma unroll
        for (int j = 0; j < KPL; ++j) af[j] = Rw[base[mt] + o[j]];
      }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = M::mma(af, bf[nt], acc[mt][nt]);
    }
  }

  // fused multiplicative update (recon:130-137)
  const int C1 = g.C1[z];
  const float* Tz = Tn + g.toffT[z];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int a2 = 16 * b2 + 4 * w + mt;
    if (a2 >= g.n2) continue;
    const int c2 = norm_class(a2, g.n2, r);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int ph = grp * 16 * NT + nt * 16 + lr;
      if (ph >= kP) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int a1 = 16 * b1 + 4 * lg + j;
        if (a1 >= g.n1) continue;
        const int c1 = norm_class(a1, g.n1, r);
        const float B1 = Tz[((size_t)c2 * C1 + c1) * kP + ph];
        const size_t idx = (((size_t)z * g.n2 + a2) * g.n1 + a1) * kP + ph;
        const float x = X[idx];
        float C = x * acc[mt][nt][j] / B1;
        if (isnan(C) || isinf(C) || C < 0.f) C = 0.f;
        float xn = C * wgt + (1.f - wgt) * x;
        if (xn < 0.f) xn = 0.f;
        X[idx] = xn;
      }
    }
  }
}

// ------------------------------------------------------------ table packing
// ABI block [e][phase] (phase fastest, 225) -> forward pack [e_pad][PHI].
template <typename T>
__global__ void k_pack_fwd(const float* __restrict__ src, T* __restrict__ dst, PackGeo pg) {
  const int z = blockIdx.y, v = blockIdx.z;
  const int K = 2 * pg.r[z] + 1, ntaps = K * K;
  const int PHI = Mma<T>::PHI;
  const size_t rows = (size_t)pg.rows[z];
  const float* s = src + (size_t)v * pg.src_view + pg.src_off[z];
  T* d = dst + (size_t)v * pg.dst_view + pg.dst_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < rows * PHI;
       o += (size_t)gridDim.x * blockDim.x) {
    // destination rows run e2-fastest (row = e1*K + e2): the 16 taps of an
    // N-tile then scatter to distinct window cells (no same-address atomics)
    const int row = (int)(o / PHI), ph = (int)(o % PHI);
    const int e1 = row / K, e2 = row - e1 * K;
    const int e = e2 * K + e1;  // ABI row order is e1-fastest
    d[o] = (T)((row < ntaps && ph < kP) ? s[(size_t)e * kP + ph] : 0.f);
  }
}

// ABI block [e][phase] -> backward pack [phase_pad = 240][e_pad] (taps contiguous).
template <typename T>
__global__ void k_pack_bwd(const float* __restrict__ src, T* __restrict__ dst, PackGeo pg) {
  const int z = blockIdx.y, v = blockIdx.z;
  const int K = 2 * pg.r[z] + 1, ntaps = K * K;
  const size_t epad = (size_t)pg.rows[z];
  const float* s = src + (size_t)v * pg.src_view + pg.src_off[z];
  T* d = dst + (size_t)v * pg.dst_view + pg.dst_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < (size_t)kPhiPadBwd * epad;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o / epad), e = (int)(o % epad);
    d[o] = (T)((e < ntaps && ph < kP) ? s[(size_t)e * kP + ph] : 0.f);
  }
}

// Normaliser table, pass 1 (one view): S[z][d2][c1][phase] = sum_{d1 in rng(c1)} Wt (FP64).
template <typename T>
__global__ void k_norm_pass1(const float* __restrict__ src, double* __restrict__ S, NormGeo ng) {
  const int z = blockIdx.y;
  const int r = ng.r[z], K = 2 * r + 1;
  const int C1 = ng.C1[z];
  const size_t total = (size_t)K * C1 * kP;
  const float* s = src + ng.src_off[z];
  double* d = S + ng.s_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o % kP);
    const int c1 = (int)((o / kP) % C1);
    const int e2 = (int)(o / ((size_t)kP * C1));
    int lo, hi;
    norm_range(c1, ng.n1, r, &lo, &hi);
    double acc = 0.0;
    for (int d1 = lo; d1 <= hi; ++d1) acc += (double)(float)(T)s[((size_t)e2 * K + (d1 + r)) * kP + ph];
    d[o] = acc;
  }
}

// pass 2: T[z][c2][c1][phase] = sum_{d2 in rng(c2)} S (rounded once to FP32).
__global__ void k_norm_pass2(const double* __restrict__ S, float* __restrict__ Tn, NormGeo ng) {
  const int z = blockIdx.y;
  const int r = ng.r[z];
  const int C1 = ng.C1[z], C2 = ng.C2[z];
  const size_t total = (size_t)C2 * C1 * kP;
  const double* s = S + ng.s_off[z];
  float* d = Tn + ng.t_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o % kP);
    const int c1 = (int)((o / kP) % C1);
    const int c2 = (int)(o / ((size_t)kP * C1));
    int lo, hi;
    norm_range(c2, ng.n2, r, &lo, &hi);
    double acc = 0.0;
    for (int d2 = lo; d2 <= hi; ++d2) acc += s[((size_t)(d2 + r) * C1 + c1) * kP + ph];
    d[o] = (float)acc;
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
size_t fwd_smem_bytes(Precision p, int rmax, int bm2) {
  const size_t b = (p == Precision::kF32) ? sizeof(float) * 2 * kFwdChunk * Mma<float>::LDB
                                          : sizeof(__bf16) * 2 * kFwdChunk * Mma<__bf16>::LDB;
  return b + sizeof(float) * 4 * (16 + 2 * rmax) * (bm2 / 4 + 2 * rmax);
}

size_t bwd_smem_bytes(Precision p, int rmax, int epad_max) {
  const int Wd = 16 + 2 * rmax;
  const size_t es = (p == Precision::kF32) ? sizeof(float) : sizeof(__bf16);
  return es * ((Wd * Wd + 7) & ~7) + sizeof(int) * epad_max;
}

hipError_t launch_fwd_mfma(Precision p, const float* X, const void* W, FwdGeo g, int nblocks,
                           size_t smem, float* partial, hipStream_t st) {
  if (p == Precision::kF32) {
    hipLaunchKernelGGL((k_fwd_mfma<float, kFwdMTF32>), dim3(nblocks), dim3(256), smem, st, X,
                       (const float*)W, g, partial);
  } else {
    hipLaunchKernelGGL((k_fwd_mfma<__bf16, kFwdMTBF16>), dim3(nblocks), dim3(256), smem, st, X,
                       (const __bf16*)W, g, partial);
  }
  return hipGetLastError();
}

hipError_t launch_ratio_gather(const float* partial, FwdGeo g, int Z, int bm2, const float* LF,
                               int u, int v, float* R, float* y_out, hipStream_t st) {
  hipLaunchKernelGGL(k_ratio_gather, dim3((g.n1 * g.n2 + 255) / 256), dim3(256), 0, st, partial, g,
                     Z, bm2, LF, u, v, R, y_out);
  return hipGetLastError();
}

hipError_t launch_bwd_mfma(Precision p, float* X, const void* Wb, const float* Tn, BwdGeo g,
                           int nblocks, size_t smem, const float* R, float w, hipStream_t st) {
  if (p == Precision::kF32) {
    hipLaunchKernelGGL(k_bwd_mfma<float>, dim3(nblocks), dim3(256), smem, st, X,
                       (const float*)Wb, Tn, g, R, w);
  } else {
    hipLaunchKernelGGL(k_bwd_mfma<__bf16>, dim3(nblocks), dim3(256), smem, st, X,
                       (const __bf16*)Wb, Tn, g, R, w);
  }
  return hipGetLastError();
}

hipError_t launch_pack_fwd(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                           hipStream_t st) {
  dim3 grid(64, Z, V);
  if (p == Precision::kF32)
    hipLaunchKernelGGL(k_pack_fwd<float>, grid, dim3(256), 0, st, src, (float*)dst, pg);
  else
    hipLaunchKernelGGL(k_pack_fwd<__bf16>, grid, dim3(256), 0, st, src, (__bf16*)dst, pg);
  return hipGetLastError();
}

hipError_t launch_pack_bwd(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                           hipStream_t st) {
  dim3 grid(64, Z, V);
  if (p == Precision::kF32)
    hipLaunchKernelGGL(k_pack_bwd<float>, grid, dim3(256), 0, st, src, (float*)dst, pg);
  else
    hipLaunchKernelGGL(k_pack_bwd<__bf16>, grid, dim3(256), 0, st, src, (__bf16*)dst, pg);
  return hipGetLastError();
}

hipError_t launch_norm_table(Precision p, const float* src_view, double* S, float* Tn_view,
                             NormGeo ng, int Z, hipStream_t st) {
  dim3 grid(128, Z);
  if (p == Precision::kF32)
    hipLaunchKernelGGL(k_norm_pass1<float>, grid, dim3(256), 0, st, src_view, S, ng);
  else
    hipLaunchKernelGGL(k_norm_pass1<__bf16>, grid, dim3(256), 0, st, src_view, S, ng);
  hipLaunchKernelGGL(k_norm_pass2, grid, dim3(256), 0, st, S, Tn_view, ng);
  return hipGetLastError();
}

}  // namespace mlf

namespace mlf {
// Allow the forward kernels more than 64 KB of dynamic LDS (gfx950 has 160 KB).
hipError_t mfma_kernel_attributes() {
  hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&k_fwd_mfma<float, kFwdMTF32>),
      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fwd_mfma<__bf16, kFwdMTBF16>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bwd_mfma<float>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bwd_mfma<__bf16>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}
}  // namespace mlf
