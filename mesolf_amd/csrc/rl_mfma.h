// Geometry descriptors and launchers of the MFMA view-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mlf {

enum class Precision { kF32 = 0, kBF16 = 1, kF16 = 2 };
constexpr int kPrecisions = 3;

constexpr int kPhiPadF32 = 228;   // 225 phases padded to the f32 MFMA K-step (4)
constexpr int kPhiPadBF16 = 256;  // 225 phases padded to the bf16 MFMA K-step (32)
constexpr int kFwdChunk = 64;     // taps per forward B chunk (4 N-tiles)
constexpr int kFwdMTF32 = 2;      // M-tiles per wave, forward, f32 (block 16 x 8 lenslets)
constexpr int kFwdMTBF16 = 4;     // M-tiles per wave, forward, bf16 (block 16 x 16 lenslets)
constexpr int kBwdNTiles = 5;     // default N-tiles per backward workgroup (80 phases; 3 also built)
constexpr int kNormRow = 256;     // floats per 1/B1 class row: 225 used, 1 KB so rows start on a 128-byte line
constexpr int kPhiPadBwd = 256;   // phase rows of a backward plane block: 3 groups of 80 phases, 5 of 48 or 4 of 64
// 16-bit backprojection variants by waves per SIMD: 2 (X of three M-tiles
// loaded before the K loop, 6-slice tap ring, copies 5 K-steps ahead) or 3
// (X loaded one M-tile ahead in the epilogue, 4-slice ring) -- fewer
// registers and LDS per workgroup, so three workgroups share a CU.
constexpr int kBwdStages = 6;
inline __host__ __device__ constexpr int bwd_stages(int occ) { return occ >= 3 ? 4 : kBwdStages; }
// Bytes before the 16-bit backprojection's shifted window: the FP32 staging
// window [Wd][Wd] and the tap ring (bwd_stages slices of [16 nt][32] taps),
// which alias (the staging window is dead once the shifted one is built).
inline __host__ __device__ int bwd_ring_bytes(int Wd, int nt, int elem_bytes, int occ) {
  const int rs = 4 * ((Wd * Wd + 3) & ~3);
  const int ring = bwd_stages(occ) * nt * 16 * 32 * elem_bytes;
  return rs > ring ? rs : ring;
}
// 16-bit backprojection tap order.  Row length K is padded to Kp8 (a multiple
// of 8) and the row count to a multiple of 4; K-step ks (32 taps) holds the
// 8-tap pieces (e2 = 4 (ks / nch) + q, e1 = 8 (ks % nch) + 0..7), q = 0..3,
// nch = Kp8 / 8.  The four pieces of a step then sit at one window column in
// four consecutive rows, so with window rows of a multiple of 16 cells the
// four 16-lane groups of an A-fragment read hit the same LDS bank groups.
inline __host__ __device__ int bwd_quad_steps(int K) { return ((K + 7) / 8) * ((K + 3) / 4); }
inline __host__ __device__ void bwd_quad_tap(int t, int K, int* e2, int* e1) {
  const int nch = (K + 7) / 8, ks = t / 32, w = t % 32, b = ks / nch;
  *e2 = 4 * b + w / 8;
  *e1 = 8 * (ks - b * nch) + w % 8;
}
// Window row stride (cells) of the 16-bit backprojection: a multiple of 16
// when the workgroup's LDS still fits (otherwise Wd, with bank conflicts).
inline __host__ __device__ int bwd_row_cells(int Wd, int nt, int occ) {
  const int a = (Wd + 15) & ~15;
  return bwd_ring_bytes(Wd, nt, 2, occ) + 16 * Wd * a <= 160 * 1024 ? a : Wd;
}
constexpr int kRatioLens = 32;    // lenslets per ratio work entry
constexpr int kRatioSlices = 8;   // plane slices summed in parallel per entry (x kRatioLens = 256 threads)
constexpr int kBwdTapPad = 32;    // FP32 backward taps padded to a multiple of 32
constexpr int kBwdTapPad16 = 64;  // 16-bit: an even number of K-steps of 32 (the K loop runs in pairs)
constexpr int kTpzMTF32 = 2;      // Toeplitz forward: default input columns per wave, f32 (1 also built)
constexpr int kTpzMTBF16 = 4;     // Toeplitz forward: default input columns per wave, bf16 (2 also built)
constexpr int kShadowRow = 256;   // BF16 volume shadow: phases per lenslet (zero beyond 225)

inline int fwd_bm2(Precision p) { return 4 * (p == Precision::kF32 ? kFwdMTF32 : kFwdMTBF16); }
// FP16 pre-scales the PSF tables (per view and plane) and the volume shadow.
inline int fwd_phi(Precision p) { return p == Precision::kF32 ? kPhiPadF32 : kPhiPadBF16; }
inline size_t elem_bytes(Precision p) { return p == Precision::kF32 ? 4 : 2; }

// One patch (reconstruction problem) as the view-step kernels see it.  A
// launch covers the work entries of one or more slots; every entry names its
// slot, so patches of different lenslet grids share one launch.
struct SlotDesc {
  float* X;               // FP32 volume [z][n2][n1][256] (phases >= 225 zero)
  void* Xb;               // 16-bit shadow [z][n2][n1][256] (BF16, or FP16 of X * xs), or null
  // BF16 Toeplitz patches: the lower plane of the split volume, or null.  With
  // it, X + 2^15 as a 32-bit pattern is split into halves: Xb holds the upper
  // half (X rounded to BF16, ties away from zero: the forward's operand), Xlo
  // the lower half.  The view loop keeps the volume in this form and the FP32
  // array is rebuilt from it (launch_merge_split) before anything else reads X.
  void* Xlo;
  const float* LF;        // measured image, column-major (15 n1) x (15 n2)
  float* R;               // ratio image [n2][n1]
  float* partial;         // forward partial windows
  const int64_t* poff;    // [Z] offset of each plane's partial windows
  const float* Tn;        // normaliser tables of view 0 for this grid's class key
  int64_t pvt;            // normaliser elements per view
  const int64_t* toffT;   // [Z] plane offset inside one view's normaliser table
  const int32_t* C1;      // [Z] normaliser classes along axis 0
  int n1, n2;
  float xs, xs_inv;       // FP16 shadow scale (1 otherwise)
};

// Work entries are processed by a grid-stride loop over *count entries, so a
// launch (or a captured graph node) of fixed grid size serves any number of
// entries written to device memory before it runs.
struct FwdGeo {
  const SlotDesc* slots;
  const int32_t* r;       // [Z] forward radius
  const int64_t* woff;    // [Z] element offset of the plane block in the packed view
  const int32_t* nchunk;  // [Z] number of 64-tap chunks (Q-scatter kernel)
  const int32_t* sched;   // [entries][4] = (slot, z, b1, b2)
  const int32_t* count;   // number of entries
  const float* finv;      // FP16: [Z] inverse PSF scale of the view (null otherwise)
};

struct RatioGeo {
  const int32_t* sched;   // [entries][2] = (slot, first lenslet); 256 lenslets per entry
  const int32_t* count;
};

struct BwdGeo {
  const SlotDesc* slots;
  const int32_t* r;       // [Z] backward radius
  const int64_t* boff;    // [Z] element offset of the plane block in the packed view
  const int32_t* epad;    // [Z] padded tap count (multiple of kBwdTapPad / kBwdTapPad16)
  const int32_t* sched;   // [entries][5] = (slot, z, b1, b2, first phase group)
  const int32_t* count;
  int view;               // selects the normaliser table of the view
  int groups;             // phase groups per entry (they share its ratio window)
};

struct PackGeo {
  const int32_t* r;        // [Z]
  const int32_t* rows;     // [Z] padded row count of the destination plane block
  const int64_t* src_off;  // [Z] offset in the ABI view block
  const int64_t* dst_off;  // [Z] offset in the packed view block
  int64_t src_view, dst_view;
  int quad;                // backward pack (16-bit types): taps in K-steps of four 8-tap pieces
                           // from four consecutive rows (bwd_quad_tap)
  const float* scale;      // [V][Z] factor applied to the taps (FP16), or null
};

struct NormGeo {
  int n1, n2;
  const int32_t* r;
  const int32_t* C1;
  const int32_t* C2;
  const int64_t* src_off;  // [Z] plane offset in the ABI (backward) view block
  const int64_t* s_off;    // [Z] FP64 scratch offsets
  const int64_t* t_off;    // [Z] normaliser table offsets within the view
  const float* scale;      // [Z] FP16 tap scale of the view, or null
};

// Normaliser classes along one axis of an n-lenslet grid for a plane of
// radius r: the backprojection of ones at lenslet a sums taps d in
// [max(-r, a-n+1), min(r, a)]; for n >= 2r+1 those ranges take 2r+1 values.
__host__ __device__ inline int norm_classes(int n, int r) { return n < 2 * r + 1 ? n : 2 * r + 1; }

__host__ __device__ inline void norm_range(int c, int n, int r, int* lo, int* hi) {
  if (n < 2 * r + 1) {
    *lo = (c - n + 1 > -r) ? c - n + 1 : -r;
    *hi = c < r ? c : r;
  } else {
    *lo = (c <= r) ? -r : c - 2 * r;
    *hi = (c < r) ? c : r;
  }
}

size_t fwd_smem_bytes(Precision p, int rmax, int bm2);
// tap-chunk buffers of the Q-scatter forward at this radius (2, or 1 when two
// do not fit in LDS beside the window)
int fwd_nbuf(Precision p, int rmax, int bm2);
size_t tpz_smem_bytes(Precision p, int rmax, int mt);
int tpz_max_radius();
// mt: input columns per wave (f32: 1 or 2; bf16: 2 or 4)
hipError_t launch_fwd_tpz(Precision p, int mt, const void* W, FwdGeo g, int grid, size_t smem,
                          hipStream_t st);
hipError_t launch_make_shadow(Precision p, const float* X, void* Xb, size_t nlens, float xs,
                              hipStream_t st);
// FP32 volume rows [nlens][256] -> split form (upper plane Xb, lower plane
// Xlo; SlotDesc::Xlo), and back.  Both are exact.
hipError_t launch_make_split(const float* X, uint16_t* Xb, uint16_t* Xlo, size_t nlens, hipStream_t st);
hipError_t launch_merge_split(const uint16_t* Xb, const uint16_t* Xlo, float* X, size_t nlens, hipStream_t st);
// scale[v][z] = 1 / max|tap| of the (view, plane) block (1 if empty); inv = max.
hipError_t launch_block_scale(const float* src, const int64_t* off, const int32_t* r,
                              int64_t per_view, int Z, int V, float* scale, float* inv,
                              hipStream_t st);
// y = sum of the forward partials; with use_lf, R = LF(u::15, v::15) ./ y
// (non-finite -> 0) into each slot's R; y_out (single-slot launches) gets y.
hipError_t launch_ratio_gather_tpz(FwdGeo g, RatioGeo rg, int grid, int Z, int cw, int u, int v,
                                   int use_lf, float* y_out, hipStream_t st);
hipError_t launch_pack_fwd_tpz(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                               hipStream_t st);
size_t bwd_smem_bytes(Precision p, int rmax, int epad_max, int nt, int occ);

hipError_t launch_fwd_mfma(Precision p, const void* W, FwdGeo g, int grid, size_t smem, int nbuf,
                           hipStream_t st);
hipError_t launch_ratio_gather(FwdGeo g, RatioGeo rg, int grid, int Z, int bm2, int u, int v,
                               int use_lf, float* y_out, hipStream_t st);
// nt: 16-phase N-tiles per workgroup (3 or 5); occ: waves per SIMD of the
// 16-bit variants (2, or 3 with nt = 3; FP32 has only 2); split (BF16 only):
// every slot holds its volume in the split form (SlotDesc::Xlo)
hipError_t launch_bwd_mfma(Precision p, int nt, int occ, bool split, const void* Wb, BwdGeo g, int grid,
                           size_t smem, float w, hipStream_t st);
hipError_t launch_pack_fwd(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                           hipStream_t st);
hipError_t launch_pack_bwd(Precision p, const float* src, void* dst, PackGeo pg, int Z, int V,
                           hipStream_t st);
hipError_t launch_norm_table(Precision p, const float* src_view, double* S, float* Tn_view,
                             NormGeo ng, int Z, hipStream_t st);
hipError_t mfma_kernel_attributes();

}  // namespace mlf
