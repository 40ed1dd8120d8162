// Handle structures shared by the host runtime modules (single-patch driver,
// batches and the serving loop).  The handle types are declared opaque in
// include/mesolf.h.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"
#include "rl_mfma.h"

struct mlf_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  mlf::DevBuf<float> mex_kq;  // MEX operator GEMM operand, reused across calls (guarded by mu)
  std::mutex mu;
  int refs = 0;         // live PSFs created on this context
  bool closed = false;  // destroy requested; freed when refs reaches 0
};

namespace mlf {

// Handle lifetimes: a PSF keeps its context alive and a patch (or batch)
// keeps its PSF alive, so destroying a parent first only defers its release.
extern std::mutex g_life_mu;
void ctx_free(mlf_context ctx);
mlf_status ctx_activate(mlf_context ctx);

template <typename T>
mlf_status upload_vec(DevBuf<T>& d, const std::vector<T>& h) {
  MLF_CHECK(d.alloc(h.size()));
  if (!h.empty()) MLF_HIP(hipMemcpy(d.ptr, h.data(), d.bytes(), hipMemcpyHostToDevice));
  return MLF_OK;
}

// PSF tables re-laid out for the MFMA kernels of one precision.
struct PrecPack {
  bool built = false;
  bool tpz = false;                   // forward tables in the Toeplitz layout [z][e1][e2pad][PHI]
  DevBuf<uint8_t> fwd, bwd;           // [V][z][e_pad64][PHI], [V][z][kPhiPadBwd][e_pad32]
  int64_t pvf = 0, pvb = 0;           // elements per view
  std::vector<int64_t> woff, boff;    // per-plane element offsets inside a view
  std::vector<int32_t> nchunk, epad;
  DevBuf<int64_t> d_woff, d_boff;
  DevBuf<int32_t> d_nchunk, d_epad;
  DevBuf<float> d_fscale, d_finv, d_bscale;  // FP16: [V][Z] tap scales (empty otherwise)
};

// Backprojection-of-ones table for one grid-size class (FP64 sums of the taps
// as the kernels see them, rounded once to FP32).
struct NormTable {
  int key1 = -1, key2 = -1;  // -1: grid at least 2*rb_max+1 lenslets on that axis
  DevBuf<float> T;
  int64_t pvt = 0;
  std::vector<int64_t> toff;
  std::vector<int32_t> C1, C2;
  DevBuf<int64_t> d_toff;
  DevBuf<int32_t> d_C1;
};

}  // namespace mlf

struct mlf_psf_s {
  mlf_context ctx = nullptr;
  int Z = 0, V = 0;
  std::vector<int32_t> uv;       // [V][2]
  std::vector<double> weight;    // [V]
  std::vector<int32_t> rf, rb;   // [Z]
  std::vector<int64_t> offf, offb;
  int64_t per_view_f = 0, per_view_b = 0;
  mlf::DevBuf<float> fwd, bwd;        // ABI layout, FP32
  mlf::DevBuf<int32_t> d_rf, d_rb;
  mlf::DevBuf<int64_t> d_offf, d_offb;
  mlf::PrecPack pack[mlf::kPrecisions];
  std::vector<std::unique_ptr<mlf::NormTable>> norms[mlf::kPrecisions];
  std::mutex mu;
  int refs = 0;         // live patches
  bool closed = false;
};

struct mlf_batch_s;

struct mlf_patch_s {
  mlf_psf psf = nullptr;
  int n1 = 0, n2 = 0, Z = 0;
  mlf_precision prec = MLF_PRECISION_FP32;
  mlf::DevBuf<float> X, LF, part, R, y, tmp;
  mlf::DevBuf<double> pb, pa, dsum;
  // filter plan
  int f_nplanes = 0, nblk = 0, f_maxlen = 1;
  mlf::DevBuf<int32_t> f_plane, f_n0_1, f_len_1, f_n0_2, f_len_2;
  mlf::DevBuf<int64_t> f_off_1, f_off_2;
  mlf::DevBuf<float> f_taps_1, f_taps_2;
  bool image_set = false, volume_set = false;
  // MFMA path
  bool reference = false;              // FP32 reference kernels instead of MFMA
  mlf::Precision kp = mlf::Precision::kF32;
  int bm2 = 8;
  bool tpz = false;                    // Toeplitz forward kernel
  int cw = 16;                         // Toeplitz: input columns per workgroup (4 * fwd_mt)
  int fwd_mt = 0, bwd_nt = 0;          // tile shape: forward columns per wave, backward N-tiles
  int bwd_grp = 0;                     // phase groups per backprojection entry (0: all of them)
  int bwd_occ = 2;                     // backprojection waves per SIMD (16-bit: 2 or 3)
  float xs = 1.f;                      // FP16: shadow = X * xs (chosen so that X * xs ~ 1)
  mlf::DevBuf<float> partial;               // forward window partials
  mlf::DevBuf<uint16_t> Xb;                 // BF16 shadow of X, [lenslet][kShadowRow] (BF16 Toeplitz)
  mlf::DevBuf<uint16_t> Xlo;                // BF16 Toeplitz: lower plane of the split volume (SlotDesc::Xlo)
  mlf::DevBuf<int64_t> d_poff;
  // work entries of this patch: host (without slot index) for batch assembly,
  // device (slot 0) for single-patch launches
  std::vector<int32_t> fent, bent;     // (z, b1, b2), (z, b1, b2, group)
  mlf::DevBuf<int32_t> d_fsched, d_bsched, d_rsched, d_counts;  // counts: fwd, bwd, ratio
  mlf::SlotDesc slot{};
  mlf::DevBuf<mlf::SlotDesc> d_slot;
  int nblk_f = 0, nblk_b = 0, nblk_r = 0;
  size_t smem_f = 0, smem_b = 0;
  int fwd_nbuf = 2;  // Q-scatter forward: tap-chunk buffers
  mlf::NormTable* norm = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  // per-kernel timing (of every timing_stride-th view step; the filter always)
  bool timing = false;
  int timing_stride = 1;
  int64_t tim_seq = 0;   // view steps since timing was enabled
  bool tim_now = false;  // the launches being enqueued are timed
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_class;  // class of the pair (2k, 2k+1)
  size_t ev_used = 0;
  mlf_batch_s* batch = nullptr;  // batch this patch is attached to
  std::mutex mu;
  ~mlf_patch_s() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  }
};

namespace mlf {
void psf_free(mlf_psf psf);  // called with g_life_mu held
mlf_status patch_filter(mlf_patch_s* p);
mlf_status patch_refresh_shadow(mlf_patch_s* p);
mlf_status patch_merge_split(mlf_patch_s* p);
// Precision of the kernels for an ABI precision.
inline Precision kernel_precision(mlf_precision p) {
  return p == MLF_PRECISION_BF16 ? Precision::kBF16
                                 : (p == MLF_PRECISION_FP16 ? Precision::kF16 : Precision::kF32);
}
void batch_mark_dirty(mlf_batch_s* b);  // a patch's slot descriptor changed
// Device bytes a patch holds (its own buffers, not the PSF-wide tables), and
// the same figure before creating one of that grid.
size_t patch_device_bytes(const mlf_patch_s* p);
size_t patch_bytes_estimate(mlf_psf_s* ps, mlf_precision prec, int n1, int n2);
mlf_status batch_forget_patch(mlf_patch_s* p);  // detach before destruction
// Workgroups go to the 8 XCDs (each with its own L2) round-robin by id.
// Reorder `width`-int entries so that each group of entries (same `group[e]`,
// groups ranked in order of first appearance) runs on one XCD: group rank g
// goes to XCD g % 8 for even rounds g / 8 and 7 - g % 8 for odd ones (a
// snake, so heavy and light groups spread evenly), and the k-th entry of
// XCD x's list lands at position x + 8 k.  Entries keep their relative order
// within an XCD.
constexpr int kXcds = 8;
void xcd_interleave(std::vector<int32_t>& ent, int width, const std::vector<int>& group);
// (Re)build a patch's work entries, partial buffers and slot descriptor for
// its tile shape (fwd_mt, bwd_nt, bwd_grp, bwd_occ); the caller synchronises the
// stream first.
mlf_status patch_setup_tiles(mlf_patch_s* p);
// A tile shape: forward input columns per wave, backprojection 16-phase
// tiles per group, phase groups per backprojection entry (0: all).
struct TileShape {
  int mt = 0, nt = 0, grp = 0, occ = 2;
  bool operator==(const TileShape& o) const {
    return mt == o.mt && nt == o.nt && grp == o.grp && occ == o.occ;
  }
};
// Candidate tile shapes of a precision, default first.
std::vector<TileShape> tile_candidates(Precision kp);
// Phase groups per backprojection entry that a shape asks for (0 -> all).
int bwd_groups_per_entry(int nt, int grp);
// Tile shape cache (per process, optionally persisted in the file named by
// MLF_TUNE_CACHE): key = precision, grid(s), PSF radii.
bool tune_cache_get(const std::string& key, Precision kp, TileShape* shape);
void tune_cache_put(const std::string& key, const TileShape& shape, double ms);
std::string tune_key(const mlf_patch_s* p, const char* scope);
}  // namespace mlf
