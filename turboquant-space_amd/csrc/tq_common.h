// tq_common.h -- device-side view of a TurboQuant space and launch helpers
// shared by the gfx950 kernels. Internal to libtq_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tq {

// The space as the kernels see it (passed by value as a kernel argument).
// Mirrors the constant members of TurboQuantSpace
// (include/turboquant/space_turbo_quant.h:207-230).
struct SpaceDev {
  int dim;          // input dim
  int D;            // padded dim: next power of two >= max(dim, 4)
  int log2D;
  int bits;         // bits per coordinate, 2..9
  int mse_bits;     // bits - 1
  int nib;          // bits <= 4: two 4-bit units per byte
  int packed;       // packed bytes per code
  int cs;           // code size = packed + 12
  int L;            // 2^(bits-1) centroids
  float scale;      // sqrt(pi/2)/sqrt(D), as the reference computes it in fp32
  float inv_sqrt_D; // 1.0f / sqrtf((float)D), correctly rounded
  // MFMA forms: the centroids enter the fp16 B operand as fp16(c * cscale)
  // plus the fp16 remainder (the coarse form of 4-bit spaces: the first term
  // alone, cscale chosen to make it exact to 2^-15); mse_unscale =
  // fl32(1 / (2^10 * cscale)) undoes the query and centroid scales in the
  // epilogue.
  float cscale;
  float mse_unscale;
  // Coarse form: the largest |c * cscale - fp16(c * cscale)| over the
  // centroids the code units can index (bounds the near pairs' correction).
  float clo_max;
  const float *cent;  // [L]   Lloyd-Max centroids (ascending)
  const float *bnd;   // [L-1] decision boundaries (ascending)
  const float *rot;   // [D]   +-1 rotation signs (rot_seed)
  const float *qjl;   // [D]   +-1 QJL signs (qjl_seed)
};

// MFMA forms: 4-bit spaces (3 MSE bits + the QJL bit) and byte-code spaces
// of padded dim >= 1024 run the coarse pass set (q_hi*c_hi, q_lo*c_hi,
// s_hi*(+-1)) with the near pairs' dropped terms added in fixed point
// (refine_flagged, tq_mfma_tile.h); 2-3-bit spaces and smaller dims run every
// lo pass. Below 1024 the passes saved cost less than the coarse epilogue's
// flags and corrections on clustered collections (a query's cluster-mates
// are all near pairs), and at padded dim 128 the dropped q_hi*c_lo term of
// byte codes moves far pairs by up to ~9e-5 relative, too close to the 1e-4
// bar.
__host__ __device__ inline bool mfma_coarse(const SpaceDev &s) {
  return s.D >= 1024 && (s.nib ? s.bits == 4 : true);
}

// fp32 trailer field at any byte alignment (4-bit codes at D = 4 are 14 bytes
// long, so their trailers are only 2-byte aligned).
__device__ inline float load_f32(const uint8_t *p, bool aligned) {
  if (aligned) return *reinterpret_cast<const float *>(p);
  uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
               ((uint32_t)p[3] << 24);
  return __uint_as_float(v);
}

// Prepared query layout: [q_rot (D) | s_q (D) | q_norm_sq | q_norm | 0 | 0].
__host__ __device__ inline long prep_stride(int D) { return 2L * D + 4; }

// ---- launchers (defined in the .hip files) ------------------------------
hipError_t launch_encode(const SpaceDev &s, const float *X, long n,
                         uint8_t *codes, hipStream_t st);
hipError_t launch_prepare(const SpaceDev &s, const float *Q, long m,
                          float *prep, hipStream_t st);
hipError_t launch_prepare_from_codes(const SpaceDev &s, const uint8_t *codes,
                                     long m, float *prep, hipStream_t st);
// rotatedCoordsBatch: sigma-normalised rotated coordinates, [n][D] floats.
hipError_t launch_rotated(const SpaceDev &s, const float *X, long n, float *out,
                          hipStream_t st);
// Plain fp32 scan: out[q*n + j] for q < m, j < n. `ws`, if given, holds
// scan_fp32_ws_bytes(s, m) bytes of scratch; without it the scan allocates
// (stream-ordered) what it needs.
hipError_t launch_scan_fp32(const SpaceDev &s, const float *prep, long m,
                            const uint8_t *codes, long n, float *out,
                            bool symmetric, hipStream_t st, void *ws = nullptr);
size_t scan_fp32_ws_bytes(const SpaceDev &s, long m);
// 4-bit codes, 32 <= D <= 1024: pair-table scan (tq_lut.hip).
bool lut_supported(const SpaceDev &s);
hipError_t launch_scan_lut(const SpaceDev &s, const float *prep, long m,
                           const uint8_t *codes, long n, float *out,
                           bool symmetric, hipStream_t st, void *ws = nullptr);
bool encode_supported(int D);
bool scan_supported(int D);

}  // namespace tq
