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
//   ones (the normaliser B1) is read from a precomputed table of 1 / B1
//   (summed and inverted in FP64).
//
// 16x16 MFMA operand layout used throughout (lane l, r = l % 16, g = l / 16):
//   A: row r, k = g*KPL + j;  B: column r, k = g*KPL + j;  C: rows 4g + j, column r.
#include <type_traits>

#include "internal.h"
#include "rl_mfma.h"

namespace mlf {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

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

// FP16 operands: same shape as BF16, 3 more mantissa bits, narrower range
// (the PSF tables are pre-scaled per view and plane, the volume shadow by a
// per-patch factor, so operands stay well inside it).
template <>
struct Mma<_Float16> {
  static constexpr int KSTEP = 32, KPL = 8;
  static constexpr int PHI = kPhiPadBF16;
  static constexpr int LDB = kPhiPadBF16;
  typedef halfx8 frag;
  __device__ static __forceinline__ floatx4 mma(frag a, frag b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ frag load(const _Float16* p) { return *(const halfx8*)p; }
};

// 16-bit shadow encoding of a value (FP32 kernels keep no shadow; BF16 stands in).
template <typename T>
__device__ __forceinline__ unsigned short to_bits16(float v) {
  if constexpr (sizeof(T) == 2) return __builtin_bit_cast(unsigned short, (T)v);
  else return __builtin_bit_cast(unsigned short, (__bf16)v);
}
template <typename T>
constexpr bool is_f16() {
  return __is_same(T, _Float16);
}

__device__ __forceinline__ int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Global-memory access through a generic pointer.  The compiler emits flat
// instructions for generic pointers it cannot prove global, and a flat
// instruction counts on the LDS counter too, so a later LDS wait or barrier
// also waits for the memory round trip.  Every pointer passed here points
// into device memory.
template <typename V>
__device__ __forceinline__ V gload(const V* p) {
  return *(const __attribute__((address_space(1))) V*)p;
}
// Read-only metadata (work entries, slot descriptors, per-plane geometry) is
// written before a launch and never by a kernel, so it is read through the
// constant address space: the compiler then uses scalar loads for uniform
// addresses.  A vector load would wait (vector-memory operations complete in
// order) for every store the wave still has in flight from its previous entry.
template <typename V>
__device__ __forceinline__ V cload(const V* p) {
  return *(const __attribute__((address_space(4))) V*)p;
}
template <typename V>
__device__ __forceinline__ void gstore(V* p, V v) {
  *(__attribute__((address_space(1))) V*)p = v;
}

// Dword 3 of a raw buffer resource on gfx9-family parts (32-bit data format,
// no swizzle, no index stride).
constexpr int kBufferDword3 = 0x00020000;

// Lane i of each 16-lane row takes lane (i - n) mod 16 of the same row (DPP
// row_ror:n), n = 0 .. 3.
__device__ __forceinline__ float row_ror(float v, int n) {
  const int i = __builtin_bit_cast(int, v);
  switch (n) {
    case 1: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(i, 0x121, 0xF, 0xF, false));
    case 2: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(i, 0x122, 0xF, 0xF, false));
    case 3: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(i, 0x123, 0xF, 0xF, false));
    default: return v;
  }
}

// Work entries and slot descriptors are read from memory the kernel also
// writes through other pointers, so the compiler cannot keep them in scalar
// registers by itself; they are wave-uniform by construction.  Values that
// feed buffer resources must be scalar, or the compiler emits a waterfall
// loop per buffer instruction.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename P>
__device__ __forceinline__ P* uni_ptr(P* p) {
  return reinterpret_cast<P*>(uni64(reinterpret_cast<int64_t>(p)));
}
// Slot descriptor with every field wave-uniform.
__device__ __forceinline__ SlotDesc load_slot(const SlotDesc* slots, int i) {
  const SlotDesc* d = slots + i;
  SlotDesc o;
  o.X = uni_ptr(cload(&d->X));
  o.Xb = uni_ptr(cload(&d->Xb));
  o.Xlo = uni_ptr(cload(&d->Xlo));
  o.LF = uni_ptr(cload(&d->LF));
  o.R = uni_ptr(cload(&d->R));
  o.partial = uni_ptr(cload(&d->partial));
  o.poff = uni_ptr(cload(&d->poff));
  o.Tn = uni_ptr(cload(&d->Tn));
  o.pvt = uni64(cload(&d->pvt));
  o.toffT = uni_ptr(cload(&d->toffT));
  o.C1 = uni_ptr(cload(&d->C1));
  o.n1 = uni(cload(&d->n1));
  o.n2 = uni(cload(&d->n2));
  o.xs = __builtin_bit_cast(float, uni(__builtin_bit_cast(int, cload(&d->xs))));
  o.xs_inv = __builtin_bit_cast(float, uni(__builtin_bit_cast(int, cload(&d->xs_inv))));
  return o;
}
// s_waitcnt vmcnt(<= n) without waiting on LDS/export counters (gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}
// Workgroup barrier that waits only for this wave's LDS operations.
// __syncthreads() is a workgroup fence plus s_barrier, and the fence waits for
// every outstanding vector-memory operation (vmcnt(0)) because LDS DMA writes
// are counted there; a caller that has already waited for the copies other
// waves will read (wait_vm_le) keeps its younger loads in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt not waited for
  __builtin_amdgcn_s_barrier();
}
__device__ __forceinline__ void wait_vm_le(int n) {
  switch (n) {
#define MLF_W(k) case k: wait_vm<k>(); break;
    MLF_W(0) MLF_W(1) MLF_W(2) MLF_W(3) MLF_W(4) MLF_W(5) MLF_W(6) MLF_W(7) MLF_W(8) MLF_W(9)
    MLF_W(10) MLF_W(11) MLF_W(12) MLF_W(13) MLF_W(14) MLF_W(15) MLF_W(16) MLF_W(17) MLF_W(18)
    MLF_W(19) MLF_W(20) MLF_W(21) MLF_W(22) MLF_W(23) MLF_W(24) MLF_W(25) MLF_W(26) MLF_W(27)
    MLF_W(28) MLF_W(29) MLF_W(30) MLF_W(31) MLF_W(32) MLF_W(33) MLF_W(34) MLF_W(35) MLF_W(36)
    MLF_W(37) MLF_W(38) MLF_W(39) MLF_W(40) MLF_W(41) MLF_W(42) MLF_W(43) MLF_W(44) MLF_W(45)
    MLF_W(46) MLF_W(47) MLF_W(48) MLF_W(49) MLF_W(50) MLF_W(51) MLF_W(52) MLF_W(53) MLF_W(54)
    MLF_W(55) MLF_W(56) MLF_W(57) MLF_W(58) MLF_W(59) MLF_W(60) MLF_W(61) MLF_W(62) MLF_W(63)
#undef MLF_W
    default: wait_vm<0>(); break;
  }
}


// ------------------------------------------------------------------ forward
// Q-scatter forward (the kernel for planes of radius > 23).  NBUF: tap chunks
// in LDS -- 2 (the next chunk is copied while the current one is used) when
// the window fits beside them, else 1 (FP32 beyond radius 22).
template <typename T, int MT, int NBUF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBUF == 1 ? 2 : 1)))
void k_fwd_mfma(const T* __restrict__ W, FwdGeo g) {
  using M = Mma<T>;
  constexpr int KSTEP = M::KSTEP, KPL = M::KPL, PHI = M::PHI, LDB = M::LDB;
  constexpr int NKS = PHI / KSTEP;
  constexpr int BM2 = 4 * MT;
  constexpr int CH = kFwdChunk;  // taps per chunk (4 N-tiles of 16)
  extern __shared__ __align__(16) char smem[];
  T* Bs = reinterpret_cast<T*>(smem);                                  // [NBUF][CH][LDB]
  float* ywin = reinterpret_cast<float*>(smem + sizeof(T) * NBUF * CH * LDB);

  const int nent = uni(cload(g.count));
  for (int ent = blockIdx.x; ent < nent; ent += gridDim.x) {
    const int* s = g.sched + 4 * ent;
    const SlotDesc sd = load_slot(g.slots, uni(cload(s)));
    const float* __restrict__ X = sd.X;
    const int n1 = sd.n1, n2 = sd.n2;
    const int z = uni(cload(s + 1)), b1 = uni(cload(s + 2)), b2 = uni(cload(s + 3));
    const int r = uni(cload(g.r + z));
    const int K = 2 * r + 1, ntaps = K * K;
    const int nch = uni(cload(g.nchunk + z));
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
        const bool ok = a1 < n1 && a2 < n2;
        const float* xp = X + (((size_t)z * n2 + (ok ? a2 : 0)) * n1 + (ok ? a1 : 0)) * kXRow;
        // branch-free: load from a clamped in-row phase, select zero afterwards
  #pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
  #pragma unroll
          for (int j = 0; j < KPL; ++j) {
            const int ph = ks * KSTEP + lg * KPL + j;
            const float v = xp[ph < kP ? ph : kP - 1];
            const float val = (ok && ph < kP) ? v : 0.f;
            if constexpr (KPL == 1) a[mt][ks] = val;
            else a[mt][ks][j] = (__bf16)val;
          }
        }
      }
    }

    const T* Wz = W + uni64(cload(g.woff + z));
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

    if (NBUF == 2) {
      issue_chunk(0, 0);
      __builtin_amdgcn_s_waitcnt(0);
    }
    for (int c = 0; c < nch; ++c) {
      __syncthreads();  // previous chunk landed / everyone done reading the buffer
      if (NBUF == 1) {
        issue_chunk(c, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
      } else if (c + 1 < nch) {
        issue_chunk(c + 1, (c + 1) & 1);
      }
      const T* Bc = Bs + (size_t)(NBUF == 2 ? (c & 1) : 0) * CH * LDB;
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
      if (NBUF == 2) __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of chunk c+1 done
    }
    __syncthreads();
    const int nb1 = ceil_div(n1, 16);
    float* out = sd.partial + uni64(cload(sd.poff + z)) + (size_t)(b2 * nb1 + b1) * W1 * W2;
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
    __syncthreads();  // LDS is reused by the next entry
  }
}

// ------------------------------------------------------- forward (Toeplitz)
// Per plane z and view:  P[(i1, c2), e2] = sum_{e1, phase} X[i1 - d1][c2][phase] * W[e1][e2][phase]
// (d1 = e1 - r), then y[i1][c2 + d2] += P.  M = 16 output rows i1 of one
// input column c2 (one M-tile per c2), N = e2 (16 * ceil(K/16)), K = (e1, phase).
// The A operand is a register window over i1: stepping e1 shifts it by one
// lane (DPP row_shr:1 within each 16-lane row) and lane 0 takes the next
// lenslet from an LDS ring filled by direct-to-LDS loads.  The B operand (one
// e1 column of taps) is copied global -> LDS while the previous column's
// MFMAs run.  The sum over e1 happens inside the MFMA accumulators, so only
// the final P tile is scattered, once per entry, by in-register diagonal sums
// and plain LDS stores.
constexpr int kTpzNT = 3;      // N-tiles (taps along axis 1): K <= 48, i.e. r <= 23
// Zero source for injection-ring rows outside the grid (read by the DMA, never
// written), so the shift needs no per-element select.
__device__ __align__(16) float g_zero_row[256] = {};  // >= one ring row (FP32: 225 floats)

// W/ring buffers: the DMA runs NB-1 steps ahead.  Two buffers keep the BF16
// kernel at 75 KB of LDS (two workgroups per CU, which measured 1.9x faster
// than three buffers at one workgroup per CU); FP32 W buffers are 45 KB each.
template <typename T>
struct TpzBufs {
  static constexpr int value = 2;
};
constexpr int kRingRow = 256;  // elements per lenslet row of the injection ring (phases >= 225 stay 0)

template <typename T, int MT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_fwd_tpz(const T* __restrict__ W, FwdGeo g) {
  using M = Mma<T>;
  constexpr int KSTEP = M::KSTEP, KPL = M::KPL, PHI = M::PHI;
  constexpr int NKS = PHI / KSTEP;
  constexpr int CW = 4 * MT;  // input columns c2 per workgroup
  constexpr int VEC = 16 / sizeof(T);
  constexpr int RV = PHI / VEC;
  constexpr bool SWZ = (RV == 32);
  constexpr int NB = TpzBufs<T>::value;                        // buffers; DMA runs NB-1 steps ahead
  constexpr int WSLOT = ((16 * kTpzNT * RV + 255) / 256) * 256;  // 16-byte slots per W buffer
  extern __shared__ __align__(16) char smem[];
  T* Wb = reinterpret_cast<T*>(smem);                                        // [NB][WSLOT*VEC]
  // injection ring: FP32 volume rows (FP32) or BF16 shadow rows (BF16), kRingRow elements each
  T* ring = reinterpret_cast<T*>(smem + (size_t)NB * WSLOT * 16);            // [NB][CW][kRingRow]
  float* ywin = reinterpret_cast<float*>(ring + NB * CW * kRingRow);         // [4 waves][16][MT + 2r]

  const int nent = uni(cload(g.count));
  for (int ent = blockIdx.x; ent < nent; ent += gridDim.x) {
    const int* s = g.sched + 4 * ent;
    const SlotDesc sd = load_slot(g.slots, uni(cload(s)));
    const float* __restrict__ X = sd.X;
    const T* __restrict__ Xb = reinterpret_cast<const T*>(sd.Xb);
    const int n1 = sd.n1, n2 = sd.n2;
    const int z = uni(cload(s + 1)), b1 = uni(cload(s + 2)), b2 = uni(cload(s + 3));
    const int r = uni(cload(g.r + z));
    const int K = 2 * r + 1;
    const int NT = (K + 15) / 16;
    const int nrow = 16 * NT;
    const int WW = MT + 2 * r;
    // w: wave index, wave-uniform (readfirstlane), so loops and bases over it run on the scalar unit
    const int tid = threadIdx.x, lane = tid & 63, w = uni(tid >> 6), lr = lane & 15, lg = lane >> 4;
    const int c2_0 = CW * b2;
    // partial-window output of the entry and (FP16) the unscaling factor, read
    // here so that their latency stays off the flush
    float* const out =
        sd.partial + uni64(cload(sd.poff + z)) + (size_t)(b2 * ceil_div(n1, 16) + b1) * 16 * (CW + 2 * r);
    float unscale = 1.f;
    if constexpr (is_f16<T>()) unscale = cload(g.finv + z) * sd.xs_inv;
    const float* Xz = X + (size_t)z * n2 * n1 * kXRow;
    const T* Xbz = KPL > 1 ? Xb + (size_t)z * n2 * n1 * kShadowRow : nullptr;  // BF16 / FP16 shadow plane
    // the shadow plane as a buffer (a plane is below 2^31 bytes, checked at patch creation)
    const __amdgpu_buffer_rsrc_t shr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(Xbz), 0, KPL > 1 ? (int)((size_t)n2 * n1 * kShadowRow * sizeof(T)) : 0, kBufferDword3);
    float* yw = ywin + w * 16 * WW;  // every cell is written by the flush
    for (int o = tid; o < NB * CW * kRingRow; o += 256) ring[o] = (T)0.f;  // pads stay zero

    // A operand at e1 = 0 (d1 = -r): lane row lr holds lenslet i1 = 16*b1 + lr + r
    typename M::frag a[MT][NKS];
  #pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int c2 = c2_0 + MT * w + mt;
      const int i1 = 16 * b1 + lr + r;
      const bool ok = c2 < n2 && i1 < n1;
      const size_t lens = ((size_t)z * n2 + (ok ? c2 : 0)) * n1 + (ok ? i1 : 0);
  #pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if constexpr (KPL == 1) {
          const int ph = ks * KSTEP + lg;
          const float v = gload(X + lens * kXRow + (ph < kP ? ph : kP - 1));
          a[mt][ks] = (ok && ph < kP) ? v : 0.f;
        } else {
          // BF16 shadow rows are 256 phases (zero beyond 225), 16-byte aligned
          const typename M::frag v =
              gload(reinterpret_cast<const typename M::frag*>(Xb + lens * kShadowRow + ks * KSTEP + lg * KPL));
          a[mt][ks] = ok ? v : typename M::frag{};
        }
      }
    }
    __syncthreads();  // ring zeroed before any DMA lands

    const T* Wz = W + uni64(cload(g.woff + z));  // [K][nrow][PHI]
    // Every wave issues the same number of DMA instructions per step (W: the
    // entry's nw_inst; ring: CW*RPASS passes / 4 waves), so a step can wait
    // for an earlier batch with a fixed vmcnt.  Only the K tap rows e2 < K
    // are copied: rows K..16*NT-1 of the buffer feed accumulator columns that
    // the flush drops, so whatever they hold does not matter.
    const int nslot = K * RV;
    const int nw_inst = (nslot + 255) / 256;
    // Copy addresses: a wave-uniform (scalar) base plus the lane's 32-bit byte
    // offset, which for the W copies does not depend on the step and is
    // computed once per entry.
    constexpr int kWInst = (16 * kTpzNT * RV + 255) / 256;  // copies per wave for the widest plane
    int woff[kWInst];
  #pragma unroll
    for (int i = 0; i < kWInst; ++i) {
      const int slot = min(256 * i + 64 * w + lane, nslot - 1);  // tail slots re-read the last vector
      const int row = (unsigned)slot / RV, pos = (unsigned)slot % RV;
      const int col = SWZ ? (pos ^ (row & 31)) : pos;
      woff[i] = (row * PHI + col * VEC) * (int)sizeof(T);
    }
    auto issue_w = [&](int t, int buf) {
      const char* src = reinterpret_cast<const char*>(Wz + (size_t)t * nrow * PHI);
      char* dst = reinterpret_cast<char*>(Wb) + (size_t)buf * WSLOT * 16;
  #pragma unroll
      for (int i = 0; i < kWInst; ++i) {
        if (i >= nw_inst) break;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + woff[i]),
                                         (__attribute__((address_space(3))) void*)(dst + (256 * i + 64 * w) * 16),
                                         16, 0, 0);
      }
    };
    // injection row needed when entering step t (t >= 1): i1 = 16*b1 + r - t
    // (rows outside the grid are loaded clamped and zeroed at the shift)
    // FP32: 225 floats per lenslet in 4 passes of 64 dwords; BF16: the 512-byte
    // shadow row in 2 passes of 64 dwords.
    constexpr int RPASS = (KPL == 1) ? 4 : 2;
    auto issue_inj = [&](int t, int slot_buf) {
      const int i1r = 16 * b1 + r - t;
      const bool rowok = i1r >= 0 && i1r < n1;
      const int i1 = min(max(i1r, 0), n1 - 1);
      char* dst = reinterpret_cast<char*>(ring + slot_buf * CW * kRingRow);
      for (int q = w; q < CW * RPASS; q += 4) {
        const int c = q / RPASS, part = q - c * RPASS;
        const bool ok = rowok && c2_0 + c < n2;
        const int c2 = min(c2_0 + c, n2 - 1);
        const size_t lens = (size_t)c2 * n1 + i1;  // within the plane
        const int d = part * 64 + lane;          // dword within the row
        if constexpr (KPL == 1) {
          if (d >= kP) continue;
        }
        const char* row = ok ? (KPL == 1 ? reinterpret_cast<const char*>(Xz + lens * kXRow)
                                         : reinterpret_cast<const char*>(Xbz + lens * kShadowRow))
                             : reinterpret_cast<const char*>(g_zero_row);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(row + 4 * d),
                                         (__attribute__((address_space(3))) void*)(dst + (c * kRingRow) * (int)sizeof(T) + 4 * part * 64),
                                         4, 0, 0);
      }
    };
    // DMA instructions per wave of a batch t >= 1 (16-bit: every W slot copy is issued)
    const int n_batch = (KPL == 1 ? nw_inst : kWInst) + CW * RPASS / 4;

    floatx4 acc[MT][kTpzNT];
  #pragma unroll
    for (int mt = 0; mt < MT; ++mt)
  #pragma unroll
      for (int nt = 0; nt < kTpzNT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};

    // batch t = W column t (+ injection row t for t >= 1); batches with index
    // >= 1 are n_batch instructions per wave
    issue_w(0, 0);
    for (int t = 1; t < NB - 1 && t < K; ++t) {
      issue_w(t, t % NB);
      issue_inj(t, t % NB);
    }
    wait_vm_le(max(0, min(K, NB - 1) - 1) * n_batch);  // batch 0 landed
    __syncthreads();
    // 16-bit: the copies of batch t + NB - 1 are spread over tile 0's K-steps
    // (one injection copy and at most one W copy per K-step, no branches), so
    // the eight waves of a CU do not queue their copies at the address unit
    // all at once.  The last step re-copies the last column and a row past the
    // grid (zeros) into the free buffer, which nothing reads.
    constexpr int kInj = CW * RPASS / 4;  // injection copies per wave and step
    static_assert(KPL == 1 || (kInj <= NKS && kWInst <= NKS), "one copy of each kind per K-step");
    for (int t = 0; t < K; ++t) {
      const int tn = min(t + NB - 1, K - 1), bn = (t + NB - 1) % NB;
      if constexpr (KPL == 1) {
        if (t + NB - 1 < K) {
          issue_w(tn, bn);
          issue_inj(tn, bn);
        }
      }
      const char* w_src = reinterpret_cast<const char*>(Wz + (size_t)tn * nrow * PHI);
      char* w_dst = reinterpret_cast<char*>(Wb) + (size_t)bn * WSLOT * 16;
      const int i1r_n = 16 * b1 + r - tn;
      const bool rowok_n = i1r_n >= 0 && i1r_n < n1;
      const int i1_n = min(max(i1r_n, 0), n1 - 1);
      char* inj_dst = reinterpret_cast<char*>(ring + bn * CW * kRingRow);
      auto issue_piece = [&](int k) __attribute__((always_inline)) {
        if constexpr (KPL > 1) {
          if (k < kWInst)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(w_src + woff[k]),
                                             (__attribute__((address_space(3))) void*)(w_dst + (256 * k + 64 * w) * 16),
                                             16, 0, 0);
          if (k < kInj) {
            // rows outside the grid: an offset past the plane, which the
            // buffer load returns as zeros (no branch, no zero-row pointer)
            const int q = w + 4 * k, c = q / RPASS, part = q - c * RPASS;
            const bool ok = rowok_n && c2_0 + c < n2;
            const int c2 = min(c2_0 + c, n2 - 1);
            const unsigned voff = ok ? (unsigned)(((c2 * n1 + i1_n) * kShadowRow + part * 128) * 2) + 4u * lane
                                     : 0x80000000u;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                shr, (__attribute__((address_space(3))) void*)(inj_dst + (c * kRingRow) * (int)sizeof(T) + 4 * part * 64),
                4, voff, 0, 0, 0);
          }
        }
      };
      const T* Wc = Wb + (size_t)(t % NB) * WSLOT * VEC;
      auto load_b = [&](int ks, int nt) {
        const int row = nt * 16 + lr;
        if constexpr (SWZ) {
          const int col = (ks * KSTEP + lg * KPL) / VEC;
          return M::load(Wc + row * PHI + ((col ^ (row & 31)) * VEC));
        } else {
          return M::load(Wc + row * PHI + ks * KSTEP + lg * KPL);
        }
      };
      // N-tile outer, K-step inner: each tile's loads and MFMAs are
      // straight-line code, and the tile count is tested twice per step
      // rather than once per group of MFMAs (a branch makes the MFMAs after
      // it wait for every LDS read issued before it)
      // B fragments are read two K-steps ahead (crossing into the next
      // N-tile's first steps), so their LDS latency hides behind the MFMAs of
      // the two steps before; reading a tile that the plane does not use is
      // harmless (the buffer holds kTpzNT tiles).
      auto load_seq = [&](int q) {  // q-th step of the tile sequence
        return load_b(q % NKS, min(q / NKS, kTpzNT - 1));
      };
      typename M::frag bq0 = load_seq(0), bq1 = load_seq(1);
      auto tile = [&](auto ntc) __attribute__((always_inline)) {
        constexpr int nt = decltype(ntc)::value;
  #pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          const typename M::frag bf = bq0;
          bq0 = bq1;
          bq1 = load_seq(nt * NKS + ks + 2);
  #pragma unroll
          for (int mt = 0; mt < MT; ++mt) acc[mt][nt] = M::mma(a[mt][ks], bf, acc[mt][nt]);
          if constexpr (nt == 0) issue_piece(ks);
        }
      };
      static_assert(kTpzNT == 3, "three N-tiles");
      tile(std::integral_constant<int, 0>{});
      if (NT > 1) tile(std::integral_constant<int, 1>{});
      if (NT > 2) tile(std::integral_constant<int, 2>{});
      // batch t+1 landed; later batches may still be in flight
      wait_vm_le(max(0, min(K, t + NB) - t - 2) * n_batch);
      __syncthreads();
      if (t + 1 < K) {
        const int i1n = 16 * b1 + r - (t + 1);
        (void)i1n;
        const T* rg = ring + ((t + 1) % NB) * CW * kRingRow;
  #pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const int c = MT * w + mt;
          // shift every 16-lane row down by one lane, then lane 0 of each row
          // alone reads the entering lenslet from the ring (the other lanes'
          // LDS reads would return data nobody keeps); rows outside the grid
          // arrived as zeros
  #pragma unroll
          for (int ks = 0; ks < NKS; ++ks) {
            if constexpr (KPL == 1) {
              a[mt][ks] = __builtin_bit_cast(
                  float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, a[mt][ks]), 0x111, 0xF, 0xF, true));
            } else {
              typedef int intx4 __attribute__((ext_vector_type(4)));
              intx4 cv = __builtin_bit_cast(intx4, a[mt][ks]);
  #pragma unroll
              for (int q = 0; q < 4; ++q) cv[q] = __builtin_amdgcn_mov_dpp(cv[q], 0x111, 0xF, 0xF, true);
              a[mt][ks] = __builtin_bit_cast(typename M::frag, cv);
            }
          }
          if (lr == 0) {
            const T* rp = rg + c * kRingRow + lg * KPL;
  #pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
              if constexpr (KPL == 1) a[mt][ks] = rp[ks * KSTEP];
              else a[mt][ks] = *reinterpret_cast<const typename M::frag*>(rp + ks * KSTEP);
            }
          }
        }
      }
    }

    // P tiles -> this wave's window: y[i1][mt + e2] = sum_mt P_mt[i1, e2] (rows
    // i1 - 16*b1, columns c2 + d2 - (c2_0 + MT*w - r)).  Lane (lr, lg) holds
    // P_mt[4 lg + j, 16 nt + lr].  Rotating tile mt right by mt lanes within
    // each 16-lane row brings the terms of column 16 nt + lr to lane lr when
    // mt <= lr; the terms that wrapped (lr < mt) belong to column
    // 16 (nt + 1) + lr and are carried into the next tile.  Columns e2 >= K
    // (padded taps) are zeroed first.  Each window cell then has one writer:
    // plain LDS stores, summed in a fixed order, and no clearing of the window.
    // FP16 undoes the PSF and volume pre-scaling here (one plane per entry).
    static_assert(MT <= 4, "row rotations of up to 3 lanes");
    floatx4 carry = {0.f, 0.f, 0.f, 0.f};
  #pragma unroll
    for (int nt = 0; nt <= kTpzNT; ++nt) {
      if (nt > NT) continue;
      floatx4 col = carry;
      carry = floatx4{0.f, 0.f, 0.f, 0.f};
      if (nt < kTpzNT && nt < NT) {
        const bool ev = nt * 16 + lr < K;
  #pragma unroll
        for (int mt = 0; mt < MT; ++mt)
  #pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float v = ev ? acc[mt][nt][j] : 0.f;
            const float x = row_ror(v, mt);
            col[j] += lr >= mt ? x : 0.f;
            carry[j] += lr >= mt ? 0.f : x;
          }
      }
      const int c = nt * 16 + lr;
      if (c < WW) {
  #pragma unroll
        for (int j = 0; j < 4; ++j) yw[(4 * lg + j) * WW + c] = is_f16<T>() ? col[j] * unscale : col[j];
      }
    }
    __syncthreads();
    // partial window stored column-major ([Wc2 columns][16 rows]): the
    // ratio gather's lanes take consecutive rows i1, so they read
    // consecutive floats
    const int Wc2 = CW + 2 * r;
    for (int o = tid; o < 16 * Wc2; o += 256) {
      const int c = o >> 4, i = o & 15;
      float sum = 0.f;
  #pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int cc = c - MT * q;
        if (cc >= 0 && cc < WW) sum += ywin[(q * 16 + i) * WW + cc];
      }
      gstore(out + o, sum);
    }
    __syncthreads();  // LDS is reused by the next entry
  }
}

// Ratio of one lenslet (recon:116-117): R = LF(u::15, v::15) ./ y with
// non-finite -> 0; y is the forward projection summed in FP64.
__device__ __forceinline__ void ratio_out(const SlotDesc& sd, int i, int i1, int i2, float yf,
                                          int u, int v, int use_lf, float* y_out) {
  if (use_lf) {
    const float lf = sd.LF[(size_t)(u + kN * i1) + (size_t)(kN * sd.n1) * (v + kN * i2)];
    float q = lf / yf;
    if (!isfinite(q)) q = 0.f;
    sd.R[i] = q;
  }
  if (y_out) y_out[i] = yf;
}

// Gather for the Toeplitz partials: y[i] = sum_z sum_{b2} window(b1 = i1/16, b2)
// in FP32, then the ratio (FP32, recon:116-117).  A workgroup takes kRatioLens lenslets; its
// kRatioSlices thread groups sum disjoint plane ranges, and the slice sums
// are added in slice order (a fixed order: the result does not depend on
// timing or on the other patches of a batch).
__global__ __launch_bounds__(256) void k_ratio_gather_tpz(FwdGeo g, RatioGeo rg, int Z, int CW, int u,
                                                          int v, int use_lf, float* __restrict__ y_out) {
  static_assert(kRatioLens * kRatioSlices == 256, "one thread per (lenslet, slice)");
  __shared__ float part[kRatioSlices][kRatioLens];
  const int li = threadIdx.x % kRatioLens, sl = threadIdx.x / kRatioLens;
  const int z0 = sl * Z / kRatioSlices, z1 = (sl + 1) * Z / kRatioSlices;
  const int nent = uni(cload(rg.count));
  for (int ent = blockIdx.x; ent < nent; ent += gridDim.x) {
    const SlotDesc sd = load_slot(g.slots, uni(cload(rg.sched + 2 * ent)));
    const int n1 = sd.n1, n2 = sd.n2;
    const int i = uni(cload(rg.sched + 2 * ent + 1)) + li;
    const bool ok = i < n1 * n2;
    const int i1 = ok ? i % n1 : 0, i2 = ok ? i / n1 : 0;
    const int nb1 = ceil_div(n1, 16), nb2 = ceil_div(n2, CW);
    const int b1 = i1 / 16;
    // The slice's planes go in chunks of kZC: the radii and offsets of a chunk
    // are read together, then up to kW windows per plane (all of them for
    // 16-column blocks, r <= 23), and only then summed, in plane then window
    // order (an unused slot adds +0.0, which changes no sum).  A plane with
    // more windows (narrower blocks) reads the rest one by one.
    float y = 0.f;
    if (ok) {
      constexpr int kZC = 8, kW = 4;
      for (int zc = z0; zc < z1; zc += kZC) {
        int rr[kZC];
        const float* pz[kZC];
  #pragma unroll
        for (int q = 0; q < kZC; ++q) {
          const int zq = min(zc + q, z1 - 1);  // the plane slice differs between the half-waves
          rr[q] = gload(g.r + zq);
          pz[q] = sd.partial + gload(sd.poff + zq);
        }
        float vv[kZC][kW];
  #pragma unroll
        for (int q = 0; q < kZC; ++q) {
          const int r = rr[q], Wc2 = CW + 2 * r;
          const int b2h = min(nb2 - 1, (i2 + r) / CW);
          const int b2l = (i2 - r - CW + 1 <= 0) ? 0 : (i2 - r - CW + 1 + CW - 1) / CW;
  #pragma unroll
          for (int wi = 0; wi < kW; ++wi) {
            const int b2 = b2l + wi;
            const bool use = zc + q < z1 && b2 <= b2h;
            const size_t idx = use ? (size_t)(b2 * nb1 + b1) * 16 * Wc2 + (i2 - CW * b2 + r) * 16 + (i1 - 16 * b1) : 0;
            const float t = gload(pz[q] + idx);
            vv[q][wi] = use ? t : 0.f;
          }
        }
  #pragma unroll
        for (int q = 0; q < kZC; ++q) {
          if (zc + q >= z1) break;
  #pragma unroll
          for (int wi = 0; wi < kW; ++wi) y += vv[q][wi];
          const int r = rr[q], Wc2 = CW + 2 * r;
          const int b2h = min(nb2 - 1, (i2 + r) / CW);
          const int b2l = (i2 - r - CW + 1 <= 0) ? 0 : (i2 - r - CW + 1 + CW - 1) / CW;
          for (int b2 = b2l + kW; b2 <= b2h; ++b2)
            y += gload(pz[q] + (size_t)(b2 * nb1 + b1) * 16 * Wc2 + (i2 - CW * b2 + r) * 16 + (i1 - 16 * b1));
        }
      }
    }
    part[sl][li] = y;
    __syncthreads();
    if (sl == 0 && ok) {
      float t = part[0][li];
  #pragma unroll
      for (int q = 1; q < kRatioSlices; ++q) t += part[q][li];
      ratio_out(sd, i, i1, i2, t, u, v, use_lf, y_out);
    }
    __syncthreads();  // part is reused by the next entry
  }
}

// ABI block [e2][e1][phase] -> Toeplitz pack [e1][e2 padded to 16*ceil(K/16)][PHI].
template <typename T>
__global__ void k_pack_fwd_tpz(const float* __restrict__ src, T* __restrict__ dst, PackGeo pg) {
  const int z = blockIdx.y, v = blockIdx.z;
  const float sc = pg.scale ? pg.scale[(size_t)v * gridDim.y + z] : 1.f;
  const int K = 2 * pg.r[z] + 1;
  const int nrow = 16 * ((K + 15) / 16);
  const int PHI = Mma<T>::PHI;
  const float* s = src + (size_t)v * pg.src_view + pg.src_off[z];
  T* d = dst + (size_t)v * pg.dst_view + pg.dst_off[z];
  const size_t total = (size_t)K * nrow * PHI;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o % PHI);
    const int e2 = (int)((o / PHI) % nrow);
    const int e1 = (int)(o / ((size_t)PHI * nrow));
    d[o] = (T)((e2 < K && ph < kP) ? s[((size_t)e2 * K + e1) * kP + ph] * sc : 0.f);
  }
}

// y[i] = sum over planes and covering blocks of the window partials (FP32,
// fixed order: plane slices as in k_ratio_gather_tpz), then the ratio
// (Q-scatter forward).
__global__ __launch_bounds__(256) void k_ratio_gather(FwdGeo g, RatioGeo rg, int Z, int BM2, int u, int v,
                                                      int use_lf, float* __restrict__ y_out) {
  __shared__ float part[kRatioSlices][kRatioLens];
  const int li = threadIdx.x % kRatioLens, sl = threadIdx.x / kRatioLens;
  const int z0 = sl * Z / kRatioSlices, z1 = (sl + 1) * Z / kRatioSlices;
  const int nent = uni(cload(rg.count));
  for (int ent = blockIdx.x; ent < nent; ent += gridDim.x) {
    const SlotDesc sd = load_slot(g.slots, uni(cload(rg.sched + 2 * ent)));
    const int n1 = sd.n1, n2 = sd.n2;
    const int i = uni(cload(rg.sched + 2 * ent + 1)) + li;
    const bool ok = i < n1 * n2;
    const int i1 = ok ? i % n1 : 0, i2 = ok ? i / n1 : 0;
    const int nb1 = ceil_div(n1, 16), nb2 = ceil_div(n2, BM2);
    float y = 0.f;
    if (ok) {
      for (int z = z0; z < z1; ++z) {
        const int r = g.r[z];  // the plane slice differs between the half-waves
        const int W1 = 16 + 2 * r, W2 = BM2 + 2 * r;
        const int b1h = min(nb1 - 1, (i1 + r) / 16);
        const int b1l = (i1 - r - 15 <= 0) ? 0 : (i1 - r - 15 + 15) / 16;
        const int b2h = min(nb2 - 1, (i2 + r) / BM2);
        const int b2l = (i2 - r - BM2 + 1 <= 0) ? 0 : (i2 - r - BM2 + 1 + BM2 - 1) / BM2;
        const float* pz = sd.partial + sd.poff[z];
        for (int b2 = b2l; b2 <= b2h; ++b2)
          for (int b1 = b1l; b1 <= b1h; ++b1)
            y += pz[(size_t)(b2 * nb1 + b1) * W1 * W2 + (i2 - BM2 * b2 + r) * W1 + (i1 - 16 * b1 + r)];
      }
    }
    part[sl][li] = y;
    __syncthreads();
    if (sl == 0 && ok) {
      float t = part[0][li];
  #pragma unroll
      for (int q = 1; q < kRatioSlices; ++q) t += part[q][li];
      ratio_out(sd, i, i1, i2, t, u, v, use_lf, y_out);
    }
    __syncthreads();  // part is reused by the next entry
  }
}

// --------------------------------------------------------------- backward
// dst[y][x] = R(a1_0 + x, a2_0 + y) for a Wd x Wd window (0 outside the
// grid), 256 threads.  The reads of up to 16 cells per thread are issued
// before any is waited for (clamped addresses, the zero applied by a select),
// so a window costs about one L2 round trip instead of one per 256 cells.
__device__ __forceinline__ void gather_window(const float* __restrict__ R, int n1, int n2, int a1_0,
                                              int a2_0, int Wd, float* __restrict__ dst) {
  constexpr int kPer = 16;
  const int tid = threadIdx.x, ncell = Wd * Wd;
  // global (not flat) loads: a pending flat load makes the compiler wait for
  // every memory operation at the next use of any loaded register
  const __attribute__((address_space(1))) float* Rg = (const __attribute__((address_space(1))) float*)R;
  for (int c0 = 0; c0 < ncell; c0 += 256 * kPer) {
    float v[kPer];
  #pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int o = min(c0 + tid + 256 * q, ncell - 1);
      const int y = o / Wd, x = o - y * Wd;
      const int a1 = a1_0 + x, a2 = a2_0 + y;
      const bool in = a1 >= 0 && a1 < n1 && a2 >= 0 && a2 < n2;
      const float t = Rg[min(max(a1, 0), n1 - 1) + n1 * min(max(a2, 0), n2 - 1)];
      v[q] = in ? t : 0.f;
    }
    // every load waited for here, on every path: a load the compiler still
    // counts as in flight later (its register reused) would make it wait at
    // the reuse for the memory operations issued since, stores included
    wait_vm<0>();
  #pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int o = c0 + tid + 256 * q;
      if (o < ncell) dst[o] = v[q];
    }
  }
}

__device__ __forceinline__ int norm_class(int a, int n, int r) {
  if (n < 2 * r + 1) return a;  // small grid: one class per lenslet
  return a < r ? a : (a > n - 1 - r ? 2 * r - (n - 1 - a) : r);
}


// SPL (BF16 only): the volume is held as its split form during the view loop
// (SlotDesc::Xlo): the epilogue reads and writes the two 16-bit planes and
// leaves the FP32 array alone.
template <typename T, int NT, int OCC, bool SPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_bwd_mfma(const T* __restrict__ Wb, BwdGeo g, float wgt) {
  using M = Mma<T>;
  constexpr int KSTEP = M::KSTEP, KPL = M::KPL;
  constexpr int MT = 4;  // NT: N-tiles (16 phases each) per workgroup
  constexpr int STG = bwd_stages(OCC);  // 16-bit tap ring depth
  // M-tiles whose X (and the first 1/B1 rows) load before the K loop; with
  // none, the epilogue loads each tile one tile ahead.  16-bit with 3 N-tiles
  // has the registers to load every tile's X and 1/B1 rows before the K loop,
  // so its epilogue waits for no load; with 4, every tile's X.
  constexpr int PRE = OCC >= 3 ? 0 : ((NT == 3 || NT == 4) && M::KPL > 1 ? MT : MT - 1);
  // 3 N-tiles: every tile's 1/B1 rows before the K loop too
  constexpr bool TPRE = PRE == MT && NT == 3;
  static_assert(!SPL || __is_same(T, __bf16), "the split volume is a BF16 form");
  // SPL N-tiles: NPR pairs, then (odd NT) one unpaired tile
  constexpr int NPR = NT / 2;
  extern __shared__ __align__(16) char smem[];
  const int nent = uni(cload(g.count));
  for (int ent = blockIdx.x; ent < nent; ent += gridDim.x) {
    const int* s = g.sched + 5 * ent;
    const SlotDesc sd = load_slot(g.slots, uni(cload(s)));
    float* __restrict__ X = sd.X;
    T* __restrict__ Xsh = reinterpret_cast<T*>(sd.Xb);
    const float* __restrict__ R = sd.R;
    const int n1 = sd.n1, n2 = sd.n2;
    // phase groups grp0 .. grp0 + groups - 1 share the entry's ratio window
    const int z = uni(cload(s + 1)), b1 = uni(cload(s + 2)), b2 = uni(cload(s + 3)), grp0 = uni(cload(s + 4));
    const int r = uni(cload(g.r + z));
    const int K = 2 * r + 1, ntaps = K * K;
    const int epad = uni(cload(g.epad + z));
    const int64_t boffz = uni64(cload(g.boff + z));
    const int Wd = 16 + 2 * r;
    // w: wave index, wave-uniform (readfirstlane), so loops and bases over it run on the scalar unit
    const int tid = threadIdx.x, lane = tid & 63, w = uni(tid >> 6), lr = lane & 15, lg = lane >> 4;
    const int nks = epad / KSTEP;

    // Fused multiplicative update (recon:130-137), set up before the K loop so
    // that the first M-tile's loads are in flight while the MFMAs run.  The
    // taps are the MFMA's A operand and the ratio window its B operand, so the
    // accumulator tile is phases x lenslets: lane (lr, lg) holds phases
    // 4lg..4lg+3 of N-tile nt for lenslet a1 = 16 b1 + lr, i.e. 16 contiguous
    // bytes of X, of the normaliser row and (8 bytes) of the BF16 shadow.
    // Loads are 16-byte buffer loads from clamped lenslets, issued one M-tile
    // ahead; only the stores are predicated.  Buffer resources span one plane
    // (a plane is below 2^31 bytes, checked at patch creation) and return
    // zero past it.
    const int C1 = uni(cload(sd.C1 + z)), C2 = norm_classes(n2, r);
    const size_t nlp = (size_t)n2 * n1;  // lenslets per plane
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc(X + z * nlp * kXRow, 0, (int)(nlp * kXRow * 4), kBufferDword3);
    const __amdgpu_buffer_rsrc_t tr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(sd.Tn + (size_t)g.view * sd.pvt + uni64(cload(sd.toffT + z))), 0, C2 * C1 * kNormRow * 4,
        kBufferDword3);
    const __amdgpu_buffer_rsrc_t sr = __builtin_amdgcn_make_buffer_rsrc(
        Xsh ? Xsh + z * nlp * kShadowRow : nullptr, 0, Xsh ? (int)(nlp * kShadowRow * 2) : 0,
        kBufferDword3);
    // SPL: the lower plane of the split volume, rows like the shadow's
    const __amdgpu_buffer_rsrc_t lor = __builtin_amdgcn_make_buffer_rsrc(
        SPL ? reinterpret_cast<uint16_t*>(sd.Xlo) + z * nlp * kShadowRow : nullptr, 0,
        SPL ? (int)(nlp * kShadowRow * 2) : 0, kBufferDword3);
    typedef float floatx4_t __attribute__((ext_vector_type(4)));
    typedef unsigned uintx4 __attribute__((ext_vector_type(4)));
    typedef unsigned uintx2 __attribute__((ext_vector_type(2)));
    const int a1 = 16 * b1 + lr;
    const bool ok1 = a1 < n1;
    const int a1c = ok1 ? a1 : n1 - 1;
    const int c1 = norm_class(a1c, n1, r);
    // window cell of this lane's lenslet in M-tile mt: base + mt * Wr (16-bit
    // windows have rows of Wr = bwd_row_cells(Wd) cells; FP32: Wd)
    const int Wr = KPL > 1 ? bwd_row_cells(Wd, NT, OCC) : Wd;
    const int base = (lr + r) + (4 * w + r) * Wr;

    // ratio window
    float* Rw = reinterpret_cast<float*>(smem);  // FP32: [Wd][Wd]
    int* toff = reinterpret_cast<int*>(smem + sizeof(float) * ((Wd * Wd + 7) & ~7));  // FP32: [epad]
    using frag_t = typename M::frag;
    // 16-bit: [Wd][Wr] after the staging window Rs / tap ring (which alias)
    frag_t* Rv = reinterpret_cast<frag_t*>(smem + bwd_ring_bytes(Wd, NT, (int)sizeof(T), OCC));
    if constexpr (KPL == 1) {
      // FP32: scalar im2col reads, taps e1-fastest
      gather_window(R, n1, n2, 16 * b1 - r, 16 * b2 - r, Wd, Rw);
      for (int e = tid; e < epad; e += 256) {
        int off = 0;  // padded taps have zero weights; any in-window read is fine
        if (e < ntaps) {
          const int e2 = e / K, e1 = e - e2 * K;
          off = -(e1 - r) - (e2 - r) * Wd;
        }
        toff[e] = off;
      }
      lds_barrier();
    } else {
      // 16-bit: a lane's 8 taps of a K-step are 8 consecutive e1 in one row
      // (bwd_quad_tap).  Cell (x, y) of the pre-shifted window holds R(x-j, y)
      // for j = 0..7, so each A fragment is a single 16-byte LDS read.
      float* Rs = reinterpret_cast<float*>(smem);  // [Wd][Wd]
      gather_window(R, n1, n2, 16 * b1 - r, 16 * b2 - r, Wd, Rs);
      lds_barrier();
      // four cells per thread per pass, their 32 reads issued together
      for (int c0 = 0; c0 < Wd * Wd; c0 += 4 * 256) {
        float rv[4][8];
  #pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int o = min(c0 + tid + 256 * q, Wd * Wd - 1);
          const int x = o % Wd;
  #pragma unroll
          for (int j = 0; j < 8; ++j) {
            const float t = Rs[max(o - j, 0)];
            rv[q][j] = x - j >= 0 ? t : 0.f;
          }
        }
  #pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int o = c0 + tid + 256 * q;
          if (o < Wd * Wd) {
            frag_t v;
  #pragma unroll
            for (int j = 0; j < 8; ++j)
              v[j] = (T)(is_f16<T>() ? fminf(rv[q][j], 65504.f) : rv[q][j]);  // FP16: clamp to the largest finite value
            const int y = o / Wd;
            Rv[o + y * (Wr - Wd)] = v;  // cell (o % Wd, y)
          }
        }
      }
      lds_barrier();  // Rv written, Rs read by every wave (the ring copies overwrite Rs)
    }

    bool pref = false;  // the group's first tap slices were issued by the previous group
    for (int gi = 0; gi < g.groups; ++gi) {  // phase group grp: 16 * NT phases from 16 * NT * grp
      const int grp = grp0 + gi;
      const T* Bz = Wb + boffz + (size_t)(grp * 16 * NT) * epad;
      floatx4 acc[MT][NT];
  #pragma unroll
      for (int mt = 0; mt < MT; ++mt)
  #pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
      // First of the 4 consecutive phases this lane holds in N-tile nt.
      // SPL: N-tiles nt and nt + 1 (nt even, < NT - 1) form a pair that
      // covers the 32 phases from a 32-aligned phase P, row j of half h being
      // phase P + 8 (j / 4) + 4 h + j % 4, so that a lane's rows in the two
      // halves are 8 consecutive phases, 16 bytes of each 16-bit plane (with
      // FP32 X, two 16-byte pieces 16 bytes apart per instruction would touch
      // twice the lines, so the other paths keep phases 16 nt + 4 lg + j);
      // N-tile NT - 1 covers 16 phases in order, after the pairs in even
      // groups and before them in odd ones (whose first phase is 16 past a
      // 32-aligned one).  The tap copies below place the rows to match.
      constexpr bool PAIR = SPL;
      const int gbase = grp * 16 * NT;
      // with an odd tile count, a group starts 16 phases past a 32-aligned
      // phase when its index is odd
      const int odd = PAIR && NT % 2 == 1 ? (grp & 1) : 0;
      auto lph = [&](int nt) -> int {
        if constexpr (PAIR) {
          if (nt < 2 * NPR) return gbase + 16 * odd + 16 * (nt & ~1) + 8 * lg + 4 * (nt & 1);
          return gbase + (odd ? 0 : 16 * (NT - 1)) + 4 * lg;
        } else {
          return gbase + 16 * nt + 4 * lg;
        }
      };
      // every phase of N-tile nt is below 225 (only the unpaired tile of the
      // last group reaches past phase 224)
      auto tfull = [&](int nt) -> bool {
        if constexpr (PAIR) {
          if (nt < 2 * NPR) return gbase + 16 * odd + 16 * (nt & ~1) + 32 <= kP;
          return gbase + (odd ? 0 : 16 * (NT - 1)) + 16 <= kP;
        } else {
          return gbase + 16 * nt + 16 <= kP;
        }
      };
      // X of M-tiles 0 .. MT-2 is loaded before the K loop (its HBM latency
      // hides under the MFMAs; the last tile's registers are taken by the
      // loop's operands and it is loaded when the loop ends); the 1/B1 rows
      // of tile 0 too, those of later tiles (L2-resident) when the tile is
      // reached
      floatx4_t xv[MT][NT], bv[NT], bva[TPRE ? MT : 1][NT];
      // SPL: dst holds raw split words, decoded where they are used so that
      // no load is waited for early: for a pair, the upper plane's 8 phases
      // in dst[nt] and the lower plane's in dst[nt + 1]; for the unpaired
      // tile, (upper pair, lower pair) of its 4 phases
      auto ldx = [&](int mt, floatx4_t* dst) {
        const int a2 = 16 * b2 + 4 * w + mt;
        const int l = (a2 < n2 ? a2 : n2 - 1) * n1 + a1c;
  #pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          if constexpr (SPL) {
            const int so = 2 * (l * kShadowRow + lph(nt));
            if (nt < 2 * NPR) {
              if (nt & 1) continue;  // loaded with its pair
              dst[nt] = __builtin_bit_cast(floatx4_t, __builtin_amdgcn_raw_buffer_load_b128(sr, so, 0, 0));
              dst[nt + 1] = __builtin_bit_cast(floatx4_t, __builtin_amdgcn_raw_buffer_load_b128(lor, so, 0, 0));
            } else {
              const uintx2 h = __builtin_amdgcn_raw_buffer_load_b64(sr, so, 0, 0);
              const uintx2 q = __builtin_amdgcn_raw_buffer_load_b64(lor, so, 0, 0);
              dst[nt] = __builtin_bit_cast(floatx4_t, uintx4{h[0], h[1], q[0], q[1]});
            }
          } else {
            dst[nt] = __builtin_bit_cast(floatx4_t,
                                         __builtin_amdgcn_raw_buffer_load_b128(xr, 4 * (l * kXRow + lph(nt)), 0, 0));
          }
        }
      };
      auto ldt = [&](int mt, floatx4_t* dst) {
        const int a2 = 16 * b2 + 4 * w + mt;
        const int to = (norm_class(a2 < n2 ? a2 : n2 - 1, n2, r) * C1 + c1) * kNormRow;
  #pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          dst[nt] = __builtin_bit_cast(floatx4_t, __builtin_amdgcn_raw_buffer_load_b128(tr, 4 * (to + lph(nt)), 0, 0));
      };
      auto load_x = [&](int mt0, int mt1) {
  #pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          if (mt >= mt0 && mt < mt1) ldx(mt, xv[mt]);
      };
      auto load_t = [&](int mt) { ldt(mt, bv); };
      // loads issued before the K loop: load_x(0, PRE) and the first (or every) 1/B1 row set
      // SPL: two loads per pair of N-tiles and two for an unpaired one
      constexpr int kEpiLoads = PRE > 0 ? PRE * (SPL ? 2 * ((NT + 1) / 2) : NT) + (TPRE ? MT * NT : NT) : 0;
      // Vector-memory operations the previous group's update issues after it
      // copied this group's first tap slices (prefetch_next), which this
      // group's first waits count instead of waiting for them (the split
      // form's update with X loaded before the K loop): every tile's stores
      // -- a prefetching group ends below phase 225, so each of its tiles
      // takes the full-tile stores, two per pair of N-tiles and two for an
      // unpaired tile -- and, with 5 N-tiles, the 1/B1 rows of tiles
      // 2 .. MT - 1 (requested after the copies).  A lower bound is safe:
      // the waits then also cover the operations not counted.
      constexpr int kUpdOps =
          (SPL && PRE > 0) ? MT * 2 * ((NT + 1) / 2) + (TPRE || NT <= 4 ? 0 : (MT - 2) * NT) : 0;
      static_assert(kEpiLoads + kUpdOps + (bwd_stages(OCC) - 2 * (OCC >= 3 ? 1 : 2)) * ((NT + 3) / 4) <= 63,
                    "wait_vm_le covers at most 63 operations in flight");

      // SPB K-steps per barrier: a step's copies are waited for, and the
      // ring refilled, once per SPB steps; the ring then runs STG - SPB
      // steps ahead.  nks is even (taps padded to 64).
      constexpr int SPB = OCC >= 3 ? 1 : 2;
      static_assert(STG >= 2 * SPB, "ring too shallow");
      // Taps go global -> LDS by DMA (global_load_lds, one copy per
      // workgroup instead of one per wave through the L1) into a ring of
      // STG K-step slices [16 NT rows][32 taps], STG - SPB steps ahead;
      // no slice past the last step is copied, and the waits count only
      // the copies issued.  The window offset of a lane's 8 taps follows the step's row
      // block and column chunk (bwd_quad_tap), stepped without divisions.
      const int nch = (K + 7) / 8;
      T* Bq = reinterpret_cast<T*>(smem);  // ring (aliases the dead staging window Rs)
      constexpr int kSlice = NT * 16 * 32;  // elements per K-step slice
      const int ncopy = w < NT ? (NT - w + 3) / 4 : 0;  // copies per step issued by this wave
      // A copy moves one N-tile's slice (16 rows x 32 taps) as 64 16-byte
      // pieces, DMA lane s landing at byte 16 s.  Piece (row, 8-tap group q)
      // goes to s = 16 q + row, so the 16 lanes reading group q at rows
      // 0..15 read 16 consecutive pieces and the four groups the same bank
      // groups (the pattern ds_read_b128 serves without conflicts).
      // SPL: slice row j of a paired N-tile (half h) is phase row
      // 8 (j / 4) + 4 h + j % 4 of the pair (lph).
      // 16-bit tables are stored tile-major (bwd_tap_index): a 16-row tile's
      // K-step is 1 KB of contiguous memory, [16 rows][32 taps].  Element
      // offset of (row R of the tile pair starting at row0, tap tq of the step):
      // (R / 16) * 16 epad + (R % 16) * 32 + tq, for R < 32.
      const int jr = lane & 15, tq = (lane >> 4) * 8;
      auto tap_off = [&](int R) { return ((R >> 4) * 16 * epad + (R & 15) * 32 + tq) * (int)sizeof(T); };
      const int tap_lane_s = tap_off(jr);                                  // unpaired tile
      const int tap_lane_p0 = tap_off(8 * (jr >> 2) + (jr & 3));           // pair, half 0
      const int tap_lane_p1 = tap_off(8 * (jr >> 2) + 4 + (jr & 3));       // pair, half 1
      const int tap_slot = (16 * lg + lr) * 8;  // element offset of (row lr, 8-tap group lg) in an N-tile slice
      // copies of K-step ks of phase group gx into ring slot `slot`: buffer
      // loads to LDS (MUBUF) over the plane's table, [kPhiPadBwd rows][epad]
      // taps.  A FLAT-encoded global_load_lds would count as an access that
      // may reach LDS out of order, and the compiler then turns every wait for
      // a load issued before it into a wait for all memory operations,
      // stores included.
      const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<T*>(Wb + boffz), 0, kPhiPadBwd * epad * (int)sizeof(T), kBufferDword3);
      auto issue_taps_g = [&](int gx, int ks, int slot) {
        if (ks >= nks) return;  // past the last step (wave-uniform)
        const int kc = ks;
        const int oddx = PAIR && NT % 2 == 1 ? (gx & 1) : 0;
        char* dst = reinterpret_cast<char*>(Bq + slot * kSlice);
        for (int nt = w; nt < NT; nt += 4) {
          const bool pr = PAIR && nt < 2 * NPR;
          // first packed row relative to the group
          const int row0 = !PAIR ? 16 * nt : pr ? 16 * oddx + 16 * (nt & ~1) : (oddx ? 0 : 16 * (NT - 1));
          const int tl = pr ? ((nt & 1) ? tap_lane_p1 : tap_lane_p0) : tap_lane_s;
          const int so = ((gx * 16 * NT + row0) * epad + kc * 512) * (int)sizeof(T);  // wave-uniform
          __builtin_amdgcn_raw_ptr_buffer_load_lds(wr, (__attribute__((address_space(3))) void*)(dst + nt * 1024),
                                                   16, tl, so, 0, 0);
        }
      };
      auto issue_taps = [&](int ks, int slot) { issue_taps_g(grp, ks, slot); };
      // the first STG - SPB slices of the next phase group of the entry,
      // issued once every wave is done with the ring, so that they land
      // while this group's update runs (16-bit)
      auto prefetch_next = [&]() {
        if constexpr (KPL > 1) {
          if (gi + 1 < g.groups) {
            lds_barrier();
  #pragma unroll
            for (int q = 0; q < STG - SPB; ++q) issue_taps_g(grp + 1, q, q);
            pref = true;
          }
        }
      };
      if constexpr (KPL == 1) {
        load_x(0, MT - 1);
        load_t(0);
        // taps in chunks of four K-steps (the pack's chunk order): a lane's
        // taps of a chunk are one 16-byte load, issued a chunk ahead
        static_assert(KSTEP == 4, "FP32 K-step of 4 taps");
        const float* Bl = Bz + (size_t)lr * epad + 4 * lg;
        floatx4_t bc[NT];
  #pragma unroll
        for (int nt = 0; nt < NT; ++nt) bc[nt] = gload(reinterpret_cast<const floatx4_t*>(Bl + (size_t)nt * 16 * epad));
        const int nc4 = nks / 4;  // epad is a multiple of 32
        for (int c = 0; c < nc4; ++c) {
          floatx4_t bx[NT];
          if (c + 1 < nc4) {
  #pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              bx[nt] = gload(reinterpret_cast<const floatx4_t*>(Bl + (size_t)nt * 16 * epad + 16 * (c + 1)));
          }
  #pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const int o = toff[(4 * c + s4) * KSTEP + lg];
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              const typename M::frag af = Rw[base + mt * Wr + o];
  #pragma unroll
              for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = M::mma(bc[nt][s4], af, acc[mt][nt]);
            }
          }
          if (c + 1 < nc4) {
  #pragma unroll
            for (int nt = 0; nt < NT; ++nt) bc[nt] = bx[nt];
          }
        }
      } else {
        // the first slices came from the previous group's prefetch: the
        // update's operations issued since are younger than them
        const int upd = pref ? kUpdOps : 0;
        if (!pref) {
  #pragma unroll
          for (int q = 0; q < STG - SPB; ++q) issue_taps(q, q);
        }
        pref = false;
        if constexpr (PRE > 0) {
          load_x(0, PRE);  // after the first slices: their waits need not cover these loads
          if constexpr (TPRE) {
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) ldt(mt, bva[mt]);
          } else {
            load_t(0);
          }
        }
        int qb = 0, qc = 0;  // row block and column chunk of step ks
        for (int ks = 0; ks < nks; ks += SPB) {
          // this wave's copies of steps ks .. ks + SPB - 1 landed: younger are
          // the copies of steps up to min(ks + STG - SPB, nks) - 1 and, while
          // those steps were copied before the loop, the epilogue loads (and
          // the previous group's update after a prefetch)
          // In the steady state (past the first slices, before the last;
          // one copy per wave and step) the count is STG - 2 SPB, waited for
          // with one test of the step: the general wait compiles to a search
          // over the 64 counts (tens of scalar branches).  Inline assembly,
          // which the compiler cannot fold back into that search.
          if (ks >= STG - SPB && ks + STG - SPB <= nks && ncopy == 1)
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(STG - 2 * SPB) : "memory");
          else
            wait_vm_le(max(0, min(ks + STG - SPB, nks) - ks - SPB) * ncopy +
                       (ks + SPB <= STG - SPB ? kEpiLoads + upd : 0));
          lds_barrier();  // everyone's copies landed; everyone is done with steps ks - SPB ..
  #pragma unroll
          for (int h = 0; h < SPB; ++h) issue_taps(ks + STG - SPB + h, (ks + STG - SPB + h) % STG);
          // one step's fragments live at a time (NT = 5 has no registers for two)
  #pragma unroll 1
          for (int h = 0; h < SPB; ++h) {
            const int kk = ks + h;
            const T* Bs = Bq + (kk % STG) * kSlice;
            // piece lg: row e2 = 4 qb + lg, taps e1 = 8 qc .. 8 qc + 7; rows
            // past K (zero weights) read row K - 1, which keeps the bank pattern
            const int e2 = min(4 * qb + lg, K - 1);
            const int o = (r - 8 * qc) - (e2 - r) * Wr;
            if (++qc == nch) {
              qc = 0;
              ++qb;
            }
            // N-tile outer: the four A fragments are read first, then each
            // N-tile's B fragment two tiles ahead of its MFMAs, so one LDS
            // round trip per step is exposed instead of one per M-tile
            // (scheduling fences keep the compiler from hoisting every B read
            // to the top, which would leave no registers for the A fragments)
            frag_t af[MT];
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) af[mt] = Rv[base + mt * Wr + o];
            frag_t bq[3];
            bq[0] = *reinterpret_cast<const frag_t*>(Bs + tap_slot);
            if (NT > 1) bq[1] = *reinterpret_cast<const frag_t*>(Bs + 512 + tap_slot);
            __builtin_amdgcn_sched_barrier(0);
  #pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
              if (nt + 2 < NT) bq[(nt + 2) % 3] = *reinterpret_cast<const frag_t*>(Bs + (nt + 2) * 512 + tap_slot);
              __builtin_amdgcn_sched_barrier(0);
  #pragma unroll
              for (int mt = 0; mt < MT; ++mt) acc[mt][nt] = M::mma(bq[nt % 3], af[mt], acc[mt][nt]);
              __builtin_amdgcn_sched_barrier(0);
            }
          }
        }
      }

      // fused multiplicative update (recon:130-137) of M-tile mt from its X
      // rows x and 1/B1 rows b.
      // Stores take no branch except in the one N-tile that reaches past phase
      // 224 (the last of 225): lanes outside the grid store at an offset past
      // the plane, which the buffer store drops (the plane is below 2^31
      // bytes, so 2^31 + a row offset is out of range and does not wrap).
      // FULL (compile time): every N-tile of the group lies below phase 225,
      // so each takes the full-tile stores without a branch, and the
      // compiler's waits for the update's loads can count those stores
      auto update_tile = [&](auto fullc, int mt, const floatx4_t* x4, const floatx4_t* b4) {
        constexpr bool FULL = decltype(fullc)::value;
        const int a2 = 16 * b2 + 4 * w + mt;
        const bool ok = ok1 && a2 < n2;
        const int l = (a2 < n2 ? a2 : n2 - 1) * n1 + a1c;
        const unsigned xbase = ok ? 4u * (unsigned)(l * kXRow) : 0x80000000u;
        const unsigned sbase = ok ? 2u * (unsigned)(l * kShadowRow) : 0x80000000u;
        uintx2 hi_prev, lo_prev;  // SPL: the first half of a pair, stored with the second
  #pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          float xo4[4];
          if constexpr (SPL) {
            // split words -> X: (upper << 16 | lower) - 2^15, per phase
            uintx4 raw;
            if (nt < 2 * NPR) {
              const uintx4 hv = __builtin_bit_cast(uintx4, x4[nt & ~1]), lv = __builtin_bit_cast(uintx4, x4[nt | 1]);
              const int h = 2 * (nt & 1);
              raw = uintx4{hv[h], hv[h + 1], lv[h], lv[h + 1]};
            } else {
              raw = __builtin_bit_cast(uintx4, x4[nt]);
            }
            xo4[0] = __builtin_bit_cast(float, ((raw[0] << 16) | (raw[2] & 0xFFFFu)) - 0x8000u);
            xo4[1] = __builtin_bit_cast(float, ((raw[0] & 0xFFFF0000u) | (raw[2] >> 16)) - 0x8000u);
            xo4[2] = __builtin_bit_cast(float, ((raw[1] << 16) | (raw[3] & 0xFFFFu)) - 0x8000u);
            xo4[3] = __builtin_bit_cast(float, ((raw[1] & 0xFFFF0000u) | (raw[3] >> 16)) - 0x8000u);
          } else {
  #pragma unroll
            for (int j = 0; j < 4; ++j) xo4[j] = x4[nt][j];
          }
          floatx4_t xn;
  #pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float x = xo4[j];
            float C = x * acc[mt][nt][j] * b4[nt][j];  // b4: 1 / B1
            // NaN, +-Inf and negatives -> 0: keep C only in the classes
            // -0, +0, +denormal, +normal (one class test and a select)
            C = __builtin_amdgcn_classf(C, 0x1E0) ? C : 0.f;
            const float v = C * wgt + (1.f - wgt) * x;
            xn[j] = v < 0.f ? 0.f : v;
          }
          const int ph = lph(nt);
          // SPL: X + 2^15 as a bit pattern, split into its upper half (X
          // rounded to BF16, ties away from zero) and its lower half
          uintx2 hi, lo;
          if constexpr (SPL) {
            const uintx4 q = __builtin_bit_cast(uintx4, xn) + 0x8000u;
            hi = uintx2{(q[0] >> 16) | (q[1] & 0xFFFF0000u), (q[2] >> 16) | (q[3] & 0xFFFF0000u)};
            lo = uintx2{(q[0] & 0xFFFFu) | (q[1] << 16), (q[2] & 0xFFFFu) | (q[3] << 16)};
          } else {
            floatx4_t sv = xn;
            if constexpr (is_f16<T>()) sv *= sd.xs;
            hi[0] = (unsigned)to_bits16<T>(sv[0]) | ((unsigned)to_bits16<T>(sv[1]) << 16);
            hi[1] = (unsigned)to_bits16<T>(sv[2]) | ((unsigned)to_bits16<T>(sv[3]) << 16);
          }
          if (FULL || tfull(nt)) {  // wave-uniform: the whole N-tile is in the row
            if constexpr (SPL) {
              if (nt < 2 * NPR) {
                if ((nt & 1) == 0) {
                  hi_prev = hi;
                  lo_prev = lo;
                  continue;
                }
                // the pair's 8 phases start 4 before this half's
                const int o = (int)(sbase + 2u * (ph - 4));
                __builtin_amdgcn_raw_buffer_store_b128(uintx4{hi_prev[0], hi_prev[1], hi[0], hi[1]}, sr, o, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(uintx4{lo_prev[0], lo_prev[1], lo[0], lo[1]}, lor, o, 0, 0);
              } else {
                __builtin_amdgcn_raw_buffer_store_b64(hi, sr, (int)(sbase + 2u * ph), 0, 0);
                __builtin_amdgcn_raw_buffer_store_b64(lo, lor, (int)(sbase + 2u * ph), 0, 0);
              }
            } else {
              __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(uintx4, xn), xr, (int)(xbase + 4u * ph), 0, 0);
              // the 16-bit paths always carry the shadow
              if constexpr (KPL > 1) __builtin_amdgcn_raw_buffer_store_b64(hi, sr, (int)(sbase + 2u * ph), 0, 0);
            }
            continue;
          }
          // the tile that reaches past phase 224
          const int xo = 4 * (l * kXRow + ph);
          const int so = 2 * (l * kShadowRow + ph);
          if constexpr (SPL) {
            // no branches: every lane issues all four stores, at an offset
            // past the plane where it has nothing to store, so that the count
            // of memory operations (which the compiler's waits follow) does
            // not depend on the lane's phases
            const unsigned o4 = ok && ph + 4 <= kP ? (unsigned)so : 0x80000000u;
            const unsigned o1 = ok && ph < kP && ph + 4 > kP ? (unsigned)so : 0x80000000u;  // phase 224
            __builtin_amdgcn_raw_buffer_store_b64(hi, sr, (int)o4, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b64(lo, lor, (int)o4, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(hi[0] & 0xFFFFu), sr, (int)o1, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(lo[0] & 0xFFFFu), lor, (int)o1, 0, 0);
          } else if (ok && ph + 4 <= kP) {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(uintx4, xn), xr, xo, 0, 0);
            if (Xsh) __builtin_amdgcn_raw_buffer_store_b64(hi, sr, so, 0, 0);
          } else if (ok && ph < kP) {  // phase 224: the last of 225
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, xn[0]), xr, xo, 0, 0);
            if (Xsh) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(hi[0] & 0xFFFFu), sr, so, 0, 0);
          }
        }
      };
      auto update = [&](auto fullc) {
        if constexpr (TPRE) {
          // every operand was loaded before the K loop
          prefetch_next();
  #pragma unroll
          for (int mt = 0; mt < MT; ++mt) update_tile(fullc, mt, xv[mt], bva[mt]);
        } else if constexpr (PRE > 0) {
          // continued from the loads issued before the K loop.
          // SPL, up to 4 N-tiles: the 1/B1 rows of every remaining tile are
          // requested at once, then the next group's first tap slices, then
          // the stores: every load the update uses is older than every store,
          // so no wait for a load covers a store (vector-memory operations
          // complete in order).  SPL, 5 N-tiles (no registers for four row
          // sets): the rows of tile mt + 1 are requested before tile mt's
          // stores, so the wait for them covers the stores of tile mt - 1
          // only.  The other paths request each tile's rows when they reach it.
          load_x(PRE, MT);
          if constexpr (SPL && NT <= 4) {
            floatx4_t bt[MT - 1][NT];
  #pragma unroll
            for (int mt = 1; mt < MT; ++mt) ldt(mt, bt[mt - 1]);
            prefetch_next();
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) update_tile(fullc, mt, xv[mt], mt == 0 ? bv : bt[mt > 0 ? mt - 1 : 0]);
          } else if constexpr (SPL) {
            floatx4_t bw[NT];
            ldt(1, bw);
            prefetch_next();  // after this group's first loads: their waits need not cover the copies
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              if (mt >= 1 && mt + 1 < MT) ldt(mt + 1, mt % 2 == 0 ? bw : bv);
              update_tile(fullc, mt, xv[mt], mt % 2 == 0 ? bv : bw);
            }
          } else {
            prefetch_next();
  #pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              if (mt > 0) load_t(mt);
              update_tile(fullc, mt, xv[mt], bv);
            }
          }
        } else {
          // X and 1/B1 rows one M-tile ahead, in two register sets
          floatx4_t xa[NT], ba[NT], xb[NT], bb[NT];
          ldx(0, xa);
          ldt(0, ba);
          prefetch_next();
  #pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            if (mt % 2 == 0) {
              if (mt + 1 < MT) {
                ldx(mt + 1, xb);
                ldt(mt + 1, bb);
              }
              update_tile(fullc, mt, xa, ba);
            } else {
              if (mt + 1 < MT) {
                ldx(mt + 1, xa);
                ldt(mt + 1, ba);
              }
              update_tile(fullc, mt, xb, bb);
            }
          }
        }
      };
      // the split form with up to 4 N-tiles only: two copies of the update
      // cost the other paths registers they do not have
      constexpr bool kTwoUpdates = SPL && NT <= 4;
      // Every X load of this group has landed (the update used them).  The
      // compiler may still count some as in flight (a path that did not read
      // a register) and then, when the next group reloads that register,
      // waits for the memory operations issued since, the update's stores
      // included.  A wait after the update tells it: with every tile full,
      // one that leaves the stores and the prefetched copies in flight (free);
      // in the last group (no copies), one that leaves the 1/B1 row loads'
      // count in flight (the X loads are older than those on every path).
      if (kTwoUpdates && gbase + 16 * NT <= kP) {
        update(std::integral_constant<bool, kTwoUpdates>{});
        if constexpr (kTwoUpdates) wait_vm<MT * 2 * ((NT + 1) / 2) + (STG - SPB) * ((NT + 3) / 4)>();
      } else {
        update(std::integral_constant<bool, false>{});
        if constexpr (kTwoUpdates) wait_vm<MT * NT>();
      }
      // LDS is reused by the next entry (the staging window aliases the
      // ring) once every wave is here.  After an entry's last group no tap
      // copy is in flight: that group prefetched none, its epilogue waited
      // for loads issued after its last copies, and vector-memory operations
      // complete in order; only stores may still be in flight, and they need
      // not be waited for.  Copies prefetched for the next group may be in
      // flight after the other groups; that group's first wait covers them.
      lds_barrier();
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

// ABI block [e][phase] -> backward pack [kPhiPadBwd phase rows][e_pad] (taps contiguous).
template <typename T>
__global__ void k_pack_bwd(const float* __restrict__ src, T* __restrict__ dst, PackGeo pg) {
  const int z = blockIdx.y, v = blockIdx.z;
  const float sc = pg.scale ? pg.scale[(size_t)v * gridDim.y + z] : 1.f;
  const int K = 2 * pg.r[z] + 1, ntaps = K * K;
  const size_t epad = (size_t)pg.rows[z];
  const float* s = src + (size_t)v * pg.src_view + pg.src_off[z];
  T* d = dst + (size_t)v * pg.dst_view + pg.dst_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < (size_t)kPhiPadBwd * epad;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o / epad), t = (int)(o % epad);
    int e2, e1;
    if (pg.quad) {
      bwd_quad_tap(t, K, &e2, &e1);
    } else {
      // FP32: taps e1-fastest, dense, stored in chunks of 16 (four K-steps of
      // 4) lane-group-major: position 16 c + 4 g + s holds tap 16 c + 4 s + g,
      // so lane group g reads its taps of the four steps with one 16-byte load
      const int lt = (t & ~15) | ((t & 3) << 2) | ((t >> 2) & 3);
      e2 = lt / K;
      e1 = lt - e2 * K;
    }
    const bool ok = ph < kP && e2 < K && e1 < K;
    // 16-bit: tile-major, [16-row tile][K-step of 32 taps][16 rows][32 taps]
    const size_t od = pg.quad ? (size_t)(ph >> 4) * 16 * epad + (size_t)(t >> 5) * 512 + (ph & 15) * 32 + (t & 31) : o;
    d[od] = (T)(ok ? s[((size_t)e2 * K + e1) * kP + ph] * sc : 0.f);
  }
}

// Normaliser table, pass 1 (one view): S[z][d2][c1][phase] = sum_{d1 in rng(c1)} Wt (FP64).
template <typename T>
__global__ void k_norm_pass1(const float* __restrict__ src, double* __restrict__ S, NormGeo ng) {
  const int z = blockIdx.y;
  const float sc = ng.scale ? ng.scale[z] : 1.f;  // the kernels' pre-scaled taps
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
    for (int d1 = lo; d1 <= hi; ++d1) acc += (double)(float)(T)(s[((size_t)e2 * K + (d1 + r)) * kP + ph] * sc);
    d[o] = acc;
  }
}

// pass 2: T[z][c2][c1][phase] = 1 / sum_{d2 in rng(c2)} S, the reciprocal taken
// in FP64 and rounded once to FP32.  A zero sum gives +Inf, so the update's
// X * B * T is non-finite (and set to 0) where B1 = 0, as X * B / B1 is.
__global__ void k_norm_pass2(const double* __restrict__ S, float* __restrict__ Tn, NormGeo ng) {
  const int z = blockIdx.y;
  const int r = ng.r[z];
  const int C1 = ng.C1[z], C2 = ng.C2[z];
  const size_t total = (size_t)C2 * C1 * kNormRow;
  const double* s = S + ng.s_off[z];
  float* d = Tn + ng.t_off[z];
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int ph = (int)(o % kNormRow);
    const int c1 = (int)((o / kNormRow) % C1);
    const int c2 = (int)(o / ((size_t)kNormRow * C1));
    if (ph >= kP) {  // row padding
      d[o] = 0.f;
      continue;
    }
    int lo, hi;
    norm_range(c2, ng.n2, r, &lo, &hi);
    double acc = 0.0;
    for (int d2 = lo; d2 <= hi; ++d2) acc += s[((size_t)(d2 + r) * C1 + c1) * kP + ph];
    d[o] = (float)(1.0 / acc);
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
static size_t fwd_smem_nbuf(Precision p, int rmax, int bm2, int nbuf) {
  const size_t b = (p == Precision::kF32) ? sizeof(float) * nbuf * kFwdChunk * Mma<float>::LDB
                                          : sizeof(__bf16) * nbuf * kFwdChunk * Mma<__bf16>::LDB;
  return b + sizeof(float) * 4 * (16 + 2 * rmax) * (bm2 / 4 + 2 * rmax);
}

int fwd_nbuf(Precision p, int rmax, int bm2) {
  return fwd_smem_nbuf(p, rmax, bm2, 2) <= 160 * 1024 ? 2 : 1;
}

size_t fwd_smem_bytes(Precision p, int rmax, int bm2) {
  return fwd_smem_nbuf(p, rmax, bm2, fwd_nbuf(p, rmax, bm2));
}

size_t bwd_smem_bytes(Precision p, int rmax, int epad_max, int nt, int occ) {
  const int Wd = 16 + 2 * rmax;
  if (p == Precision::kF32) return sizeof(float) * ((Wd * Wd + 7) & ~7) + sizeof(int) * epad_max;
  return bwd_ring_bytes(Wd, nt, 2, occ) + 16 * (size_t)Wd * bwd_row_cells(Wd, nt, occ);
}

hipError_t launch_fwd_mfma(Precision p, const void* W, FwdGeo g, int grid, size_t smem, int nbuf,
                           hipStream_t st) {
  if (p == Precision::kF32) {
    if (nbuf == 1)
      hipLaunchKernelGGL((k_fwd_mfma<float, kFwdMTF32, 1>), dim3(grid), dim3(256), smem, st,
                         (const float*)W, g);
    else
      hipLaunchKernelGGL((k_fwd_mfma<float, kFwdMTF32, 2>), dim3(grid), dim3(256), smem, st,
                         (const float*)W, g);
  } else if (nbuf == 1) {
    hipLaunchKernelGGL((k_fwd_mfma<__bf16, kFwdMTBF16, 1>), dim3(grid), dim3(256), smem, st,
                       (const __bf16*)W, g);
  } else {
    hipLaunchKernelGGL((k_fwd_mfma<__bf16, kFwdMTBF16, 2>), dim3(grid), dim3(256), smem, st,
                       (const __bf16*)W, g);
  }
  return hipGetLastError();
}

size_t tpz_smem_bytes(Precision p, int rmax, int MT) {
  const bool f32 = p == Precision::kF32;
  const int RV = f32 ? Mma<float>::PHI / 4 : Mma<__bf16>::PHI / 8;
  const int NB = f32 ? TpzBufs<float>::value : TpzBufs<__bf16>::value;
  const size_t wslot = ((16 * kTpzNT * RV + 255) / 256) * 256;
  return (size_t)NB * wslot * 16 + elem_bytes(p) * NB * 4 * MT * kRingRow +
         sizeof(float) * 4 * 16 * (MT + 2 * rmax);
}

int tpz_max_radius() { return (16 * kTpzNT - 1) / 2; }

hipError_t launch_fwd_tpz(Precision p, int mt, const void* W, FwdGeo g, int grid, size_t smem,
                          hipStream_t st) {
  if (p == Precision::kF32) {
    if (mt == 1)
      hipLaunchKernelGGL((k_fwd_tpz<float, 1>), dim3(grid), dim3(256), smem, st, (const float*)W, g);
    else
      hipLaunchKernelGGL((k_fwd_tpz<float, 2>), dim3(grid), dim3(256), smem, st, (const float*)W, g);
  } else if (p == Precision::kF16) {
    if (mt == 2)
      hipLaunchKernelGGL((k_fwd_tpz<_Float16, 2>), dim3(grid), dim3(256), smem, st, (const _Float16*)W, g);
    else
      hipLaunchKernelGGL((k_fwd_tpz<_Float16, 4>), dim3(grid), dim3(256), smem, st, (const _Float16*)W, g);
  } else {
    if (mt == 2)
      hipLaunchKernelGGL((k_fwd_tpz<__bf16, 2>), dim3(grid), dim3(256), smem, st, (const __bf16*)W, g);
    else
      hipLaunchKernelGGL((k_fwd_tpz<__bf16, 4>), dim3(grid), dim3(256), smem, st, (const __bf16*)W, g);
  }
  return hipGetLastError();
}

template <typename T>
__global__ void k_make_shadow(const float* __restrict__ X, T* __restrict__ Xb, size_t nlens, float xs) {
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < nlens * kShadowRow;
       o += (size_t)gridDim.x * blockDim.x) {
    const size_t l = o / kShadowRow;
    const int ph = (int)(o - l * kShadowRow);
    const float v = ph < kP ? X[l * kXRow + ph] : 0.f;
    Xb[o] = (T)(is_f16<T>() ? v * xs : v);
  }
}

hipError_t launch_make_shadow(Precision p, const float* X, void* Xb, size_t nlens, float xs,
                              hipStream_t st) {
  size_t g = (nlens * kShadowRow + 255) / 256;
  if (g > 65536) g = 65536;
  if (p == Precision::kF16)
    hipLaunchKernelGGL(k_make_shadow<_Float16>, dim3((unsigned)g), dim3(256), 0, st, X,
                       (_Float16*)Xb, nlens, xs);
  else
    hipLaunchKernelGGL(k_make_shadow<__bf16>, dim3((unsigned)g), dim3(256), 0, st, X, (__bf16*)Xb,
                       nlens, xs);
  return hipGetLastError();
}

// Split form of the FP32 volume (SlotDesc::Xlo): p = bits(X) + 2^15, upper
// half to Xb, lower half to Xlo; X = bits((upper << 16 | lower) - 2^15).  Four
// phases per thread; rows are 256 phases in all three arrays.
__global__ void k_make_split(const float* __restrict__ X, uint16_t* __restrict__ Xb,
                             uint16_t* __restrict__ Xlo, size_t nq) {
  typedef unsigned uintx4 __attribute__((ext_vector_type(4)));
  typedef unsigned uintx2 __attribute__((ext_vector_type(2)));
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nq; i += (size_t)gridDim.x * blockDim.x) {
    const uintx4 q = reinterpret_cast<const uintx4*>(X)[i] + 0x8000u;
    reinterpret_cast<uintx2*>(Xb)[i] =
        uintx2{(q[0] >> 16) | (q[1] & 0xFFFF0000u), (q[2] >> 16) | (q[3] & 0xFFFF0000u)};
    reinterpret_cast<uintx2*>(Xlo)[i] = uintx2{(q[0] & 0xFFFFu) | (q[1] << 16), (q[2] & 0xFFFFu) | (q[3] << 16)};
  }
}

__global__ void k_merge_split(const uint16_t* __restrict__ Xb, const uint16_t* __restrict__ Xlo,
                              float* __restrict__ X, size_t nq) {
  typedef unsigned uintx4 __attribute__((ext_vector_type(4)));
  typedef unsigned uintx2 __attribute__((ext_vector_type(2)));
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nq; i += (size_t)gridDim.x * blockDim.x) {
    const uintx2 h = reinterpret_cast<const uintx2*>(Xb)[i];
    const uintx2 l = reinterpret_cast<const uintx2*>(Xlo)[i];
    const uintx4 p{(h[0] << 16) | (l[0] & 0xFFFFu), (h[0] & 0xFFFF0000u) | (l[0] >> 16),
                   (h[1] << 16) | (l[1] & 0xFFFFu), (h[1] & 0xFFFF0000u) | (l[1] >> 16)};
    reinterpret_cast<uintx4*>(X)[i] = p - 0x8000u;
  }
}

static unsigned stream_grid(size_t n) {
  const size_t g = (n + 255) / 256;
  return (unsigned)(g > 8192 ? 8192 : (g ? g : 1));
}

hipError_t launch_make_split(const float* X, uint16_t* Xb, uint16_t* Xlo, size_t nlens, hipStream_t st) {
  const size_t nq = nlens * (kXRow / 4);
  static_assert(kXRow == kShadowRow, "the split planes share the volume's row length");
  hipLaunchKernelGGL(k_make_split, dim3(stream_grid(nq)), dim3(256), 0, st, X, Xb, Xlo, nq);
  return hipGetLastError();
}

hipError_t launch_merge_split(const uint16_t* Xb, const uint16_t* Xlo, float* X, size_t nlens, hipStream_t st) {
  const size_t nq = nlens * (kXRow / 4);
  hipLaunchKernelGGL(k_merge_split, dim3(stream_grid(nq)), dim3(256), 0, st, Xb, Xlo, X, nq);
  return hipGetLastError();
}

// Largest |tap| of each (view, plane) block of an ABI table -> 1 / max (1 for
// an empty block): the FP16 pre-scaling factors.
__global__ void k_block_scale(const float* __restrict__ src, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ r, int64_t per_view, float* __restrict__ scale,
                              float* __restrict__ inv) {
  const int z = blockIdx.x, v = blockIdx.y, Z = gridDim.x;
  const int K = 2 * r[z] + 1;
  const size_t n = (size_t)K * K * kP;
  const float* s = src + (size_t)v * per_view + off[z];
  float m = 0.f;
  for (size_t o = threadIdx.x; o < n; o += blockDim.x) m = fmaxf(m, fabsf(s[o]));
  for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_down(m, d, 64));
  __shared__ float sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float mx = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    scale[(size_t)v * Z + z] = mx > 0.f ? 1.f / mx : 1.f;
    if (inv) inv[(size_t)v * Z + z] = mx > 0.f ? mx : 1.f;
  }
}

hipError_t launch_block_scale(const float* src, const int64_t* off, const int32_t* r,
                              int64_t per_view, int Z, int V, float* scale, float* inv,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_block_scale, dim3(Z, V), dim3(256), 0, st, src, off, r, per_view, scale, inv);
  return hipGetLastError();
}

hipError_t launch_ratio_gather_tpz(FwdGeo g, RatioGeo rg, int grid, int Z, int cw, int u, int v,
                                   int use_lf, float* y_out, hipStream_t st) {
  hipLaunchKernelGGL(k_ratio_gather_tpz, dim3(grid), dim3(256), 0, st, g, rg, Z, cw, u, v, use_lf,
                     y_out);
  return hipGetLastError();
}

hipError_t launch_pack_fwd_tpz(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                               hipStream_t st) {
  dim3 grid(64, Z, V);
  if (p == Precision::kF32)
    hipLaunchKernelGGL(k_pack_fwd_tpz<float>, grid, dim3(256), 0, st, src, (float*)dst, pg);
  else if (p == Precision::kF16)
    hipLaunchKernelGGL(k_pack_fwd_tpz<_Float16>, grid, dim3(256), 0, st, src, (_Float16*)dst, pg);
  else
    hipLaunchKernelGGL(k_pack_fwd_tpz<__bf16>, grid, dim3(256), 0, st, src, (__bf16*)dst, pg);
  return hipGetLastError();
}

hipError_t launch_ratio_gather(FwdGeo g, RatioGeo rg, int grid, int Z, int bm2, int u, int v,
                               int use_lf, float* y_out, hipStream_t st) {
  hipLaunchKernelGGL(k_ratio_gather, dim3(grid), dim3(256), 0, st, g, rg, Z, bm2, u, v, use_lf,
                     y_out);
  return hipGetLastError();
}

template <typename T, bool SPL>
static void launch_bwd_t(int nt, int occ, const void* Wb, BwdGeo g, int grid, size_t smem, float w,
                         hipStream_t st) {
  const T* W = (const T*)Wb;
  // three waves per SIMD only with 3 N-tiles (with 5 the accumulators and
  // operands exceed 168 registers)
  if constexpr (__is_same(T, __bf16)) {
    // BF16 only: 4 N-tiles, phase groups of 64 (128 bytes of each 16-bit plane)
    if (nt == 4) {
      hipLaunchKernelGGL((k_bwd_mfma<T, 4, 2, SPL>), dim3(grid), dim3(256), smem, st, W, g, w);
      return;
    }
  }
  if (nt == 3 && occ >= 3)
    hipLaunchKernelGGL((k_bwd_mfma<T, 3, 3, SPL>), dim3(grid), dim3(256), smem, st, W, g, w);
  else if (nt == 3)
    hipLaunchKernelGGL((k_bwd_mfma<T, 3, 2, SPL>), dim3(grid), dim3(256), smem, st, W, g, w);
  else
    hipLaunchKernelGGL((k_bwd_mfma<T, 5, 2, SPL>), dim3(grid), dim3(256), smem, st, W, g, w);
}

hipError_t launch_bwd_mfma(Precision p, int nt, int occ, bool split, const void* Wb, BwdGeo g, int grid,
                           size_t smem, float w, hipStream_t st) {
  if (split && p != Precision::kBF16) return hipErrorInvalidValue;
  if (p == Precision::kF32) {
    if (nt == 3)
      hipLaunchKernelGGL((k_bwd_mfma<float, 3, 2, false>), dim3(grid), dim3(256), smem, st, (const float*)Wb, g, w);
    else
      hipLaunchKernelGGL((k_bwd_mfma<float, 5, 2, false>), dim3(grid), dim3(256), smem, st, (const float*)Wb, g, w);
  } else if (p == Precision::kF16) {
    launch_bwd_t<_Float16, false>(nt, occ, Wb, g, grid, smem, w, st);
  } else if (split) {
    launch_bwd_t<__bf16, true>(nt, occ, Wb, g, grid, smem, w, st);
  } else {
    launch_bwd_t<__bf16, false>(nt, occ, Wb, g, grid, smem, w, st);
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
  else if (p == Precision::kF16)
    hipLaunchKernelGGL(k_pack_bwd<_Float16>, grid, dim3(256), 0, st, src, (_Float16*)dst, pg);
  else
    hipLaunchKernelGGL(k_pack_bwd<__bf16>, grid, dim3(256), 0, st, src, (__bf16*)dst, pg);
  return hipGetLastError();
}

hipError_t launch_norm_table(Precision p, const float* src_view, double* S, float* Tn_view,
                             NormGeo ng, int Z, hipStream_t st) {
  dim3 grid(128, Z);
  if (p == Precision::kF32)
    hipLaunchKernelGGL(k_norm_pass1<float>, grid, dim3(256), 0, st, src_view, S, ng);
  else if (p == Precision::kF16)
    hipLaunchKernelGGL(k_norm_pass1<_Float16>, grid, dim3(256), 0, st, src_view, S, ng);
  else
    hipLaunchKernelGGL(k_norm_pass1<__bf16>, grid, dim3(256), 0, st, src_view, S, ng);
  hipLaunchKernelGGL(k_norm_pass2, grid, dim3(256), 0, st, S, Tn_view, ng);
  return hipGetLastError();
}

}  // namespace mlf

namespace mlf {
// Allow the forward kernels more than 64 KB of dynamic LDS (gfx950 has 160 KB).
hipError_t mfma_kernel_attributes() {
  const void* kernels[] = {
      reinterpret_cast<const void*>(&k_fwd_mfma<float, kFwdMTF32, 1>),
      reinterpret_cast<const void*>(&k_fwd_mfma<float, kFwdMTF32, 2>),
      reinterpret_cast<const void*>(&k_fwd_mfma<__bf16, kFwdMTBF16, 1>),
      reinterpret_cast<const void*>(&k_fwd_mfma<__bf16, kFwdMTBF16, 2>),
      reinterpret_cast<const void*>(&k_fwd_tpz<float, 1>),
      reinterpret_cast<const void*>(&k_fwd_tpz<float, 2>),
      reinterpret_cast<const void*>(&k_fwd_tpz<__bf16, 2>),
      reinterpret_cast<const void*>(&k_fwd_tpz<__bf16, 4>),
      reinterpret_cast<const void*>(&k_bwd_mfma<float, 3, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<float, 5, 2, false>),
      reinterpret_cast<const void*>(&k_fwd_tpz<_Float16, 2>),
      reinterpret_cast<const void*>(&k_fwd_tpz<_Float16, 4>),
      reinterpret_cast<const void*>(&k_bwd_mfma<_Float16, 3, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<_Float16, 5, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<_Float16, 3, 3, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 3, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 5, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 3, 3, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 3, 2, true>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 5, 2, true>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 3, 3, true>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 4, 2, false>),
      reinterpret_cast<const void*>(&k_bwd_mfma<__bf16, 4, 2, true>),
  };
  for (const void* k : kernels) {
    const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace mlf
