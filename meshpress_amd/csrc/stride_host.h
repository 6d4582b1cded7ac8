// Host-side (CPU) part of the STRIDE codec: encoder front-end (meshlet
// partition, per-meshlet plan, strips, walk) and the DPR4 v3 bit writer.
//
// Internal header shared by stride_host.cpp and the C-ABI glue.
#pragma once

#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

namespace stride {

// DPR4 v3 container constants (encoder/paradelta_v5_dup.py:76-77).
constexpr uint32_t kMagicDup   = 0x44505234u;
constexpr uint32_t kVersionDup = 3u;
constexpr int      kHeaderBits = 688;   // 86 bytes, byte aligned
constexpr int      kFifoSize   = 16;    // REUSE_BUF_SIZE (encoder/paradelta_codec.py:104)
constexpr int      kReuseBits  = 5;     // ceil(log2(16 + 1))

inline int idx_bits_for(int n_local) {          // encoder/paradelta_codec.py:124
  int b = 1;
  while ((1 << b) < n_local + 1) ++b;
  return b;
}

// One meshlet after planning.  Local vertex ids are < n_local <= 256.
struct Plan {
  std::vector<int64_t> l2g;          // local -> global vertex id
  std::vector<uint8_t> tris_local;   // n_tris x 3 local ids, partition order
  std::vector<int32_t> strip_len;    // per strip
  std::vector<int32_t> strip_tris;   // concatenated local tri indices
  int n_bnd = 0;
};

// Walk entry (encoder/paradelta_v5_dup.py:114-152): kind 0 = strip-root
// vertex (anchor or delta), kind 1 = parallelogram vertex with refs a,b,c.
struct WalkEntry {
  uint8_t v, kind, a, b, c;
};

struct Prep {
  int64_t n_v = 0, n_t = 0;
  std::vector<Plan> plans;
  std::vector<std::vector<WalkEntry>> walks;
  // global ids of the vertices shared by >= 2 meshlets, in boundary-table
  // order (sorted, then stable Morton order); plan boundary locals are
  // sorted by their position here
  std::vector<int64_t> boundary;
  std::string error;
};

// Front-end parameters fixed by the caller (Python host computes the
// floating-point constants with numpy so they match the reference exactly).
struct PrepParams {
  int max_verts = 256;
  int max_tris = 256;
  double K[6];      // LEARNED_WEIGHTS[i] * inv_std[i]
  double NM[6];     // -mean[i]
  // Strip seeds drawn by np.random.default_rng(0).choice(n_f, min(6, n_f-2),
  // replace=False), one row of 6 per n_f in [0, seed_rows).
  const int32_t* seed_table = nullptr;
  int seed_rows = 0;
  int n_threads = 1;
};

// Builds partition + plans + walks.  tris: n_t x 3 int64; vn, fn: float64
// normalised positions and face normals; codes: n_v x 3 int64 global codes.
bool build_prep(const int64_t* tris, int64_t n_t, const double* vn, int64_t n_v,
                const double* fn, const int64_t* codes, const PrepParams& P,
                Prep& out);

// CPython set iteration order for a set of small non-negative ints built by
// successive add() calls in the given order.
std::vector<int> pyset_order(const std::vector<int>& inserted);

// MSB-first bit writer (utils/bit_codec.py:16-189).
class BitWriter {
 public:
  void put(uint64_t value, int n);           // n <= 64
  void put_rice(uint64_t u, int k);
  void put_eg(uint64_t u, int k);
  uint64_t bit_pos() const { return nbits_; }
  std::vector<uint8_t> finish();
  std::vector<uint8_t>& bytes() { return buf_; }
 private:
  std::vector<uint8_t> buf_;
  uint64_t acc_ = 0;   // pending bits, right aligned
  int nacc_ = 0;       // 0..7 pending bits after each put
  uint64_t nbits_ = 0;
};

struct Header {
  float center[3], scale, per_coord_err, g_min[3], g_range[3];
  int g_bits[3];
  int32_t pred_n[9];
  int32_t pred_K[3];
};

// Residual layout per meshlet m: 3 axes x n_delta[m] delta residuals
// then 3 axes x n_para[m] parallelogram residuals, axis-major.
struct Residuals {
  std::vector<int64_t> anchor;     // 3 per meshlet
  std::vector<int64_t> off;        // n_m + 1 offsets into vals (units: values)
  std::vector<int32_t> n_delta, n_para;
  std::vector<int64_t> vals;
  int64_t max_abs = 0;
};

void compute_residuals(const Prep& P, const int64_t* codes, const int32_t* pred_n,
                       const int32_t* pred_K, Residuals& R, int n_threads);

// Per-meshlet features of the learned proposal, kFeatures floats per
// selected meshlet (all when sel is null; layout in include/stride_hip.h,
// stride_encoder_features); stream statistics from R, or from canonical
// residuals when R is null.
constexpr int kFeatures = 32;
void meshlet_features(const Prep& P, const int64_t* codes, const Residuals* R, const int32_t* sel,
                      int64_t n_sel, float* out, int n_threads);

// Parallelogram samples of meshlets sel[0..n_sel) (all when sel is null):
// off gets the row offsets; cols (when not null) the codes of a, b, c, v
// as cols[(q * 3 + d) * n + i], n = off[rows].
void para_columns(const Prep& P, const int64_t* codes, const int32_t* sel, int64_t n_sel,
                  int32_t* cols, int64_t* off, int n_threads);


// DPRB v4 side-table blob <-> DPR4 v3 (stride_dprb.cpp).  Return 0 or a
// negative status (the STRIDE_E_* values) with a message in err.
struct DprbTables {
  uint32_t n_m = 0, n_v = 0, n_t = 0;
  uint64_t body_start_bit = 0;                  // of the rebuilt DPR4 stream
  std::vector<uint64_t> rel_off, resid_off;     // body-relative bit offsets
  std::vector<uint32_t> n_local, n_tris, n_strips, v_off, t_off, n_kind0;
  std::vector<uint16_t> subs;                   // 5 per meshlet
};
int dprb_from_dpr4(const uint8_t* data, size_t len, std::vector<uint8_t>& out, std::string& err,
                   int& code);
int dprb_to_dpr4(const uint8_t* blob, size_t len, std::vector<uint8_t>& dpr4, DprbTables& t,
                 std::string& err, int& code);

// PDLT v5 writer (encoder/paradelta_v5.py:144-283 and
// encoder/paradelta_v5_nb.py:20-539): LIN3 predictor with a float32
// closed-loop chain, global boundary table, Rice-coded boundary refs.
struct V5Header {
  float center[3], scale, per_coord_err, g_min[3], g_range[3];
  int g_bits[3];
};
// Fit samples of the plain-parallelogram pass (w3 == nullptr) as rows of
// (a, b, c, true) x 3 doubles, or the residual codes (n_int x 3 per
// meshlet, meshlet order) of the LIN3 pass.
void v5_samples(const Prep& P, const int64_t* codes, const double* vn, const V5Header& H,
                std::vector<double>& rows, int n_threads);
std::vector<uint8_t> write_v5(const Prep& P, const int64_t* codes, const double* vn,
                              const V5Header& H, const float* lin3, int n_threads);

// SBLB v1 blob (PDLT v5 stream + the reference bench's tables) <-> PDLT v5
// (pdlt_sblb.cpp).  Return 0 or a negative status with a message in err.
int sblb_from_pdlt(const uint8_t* data, size_t len, std::vector<uint8_t>& out, std::string& err);
// The embedded stream of a blob whose size matches its header.
int sblb_split(const uint8_t* blob, size_t len, const uint8_t*& stream, size_t& stream_len,
               std::string& err);
// sblb_split plus: every table equals what the pre-pass derives from the stream.
int sblb_check(const uint8_t* blob, size_t len, std::string& err);

// tagk: per meshlet 12 bytes: (tag,k) x 3 delta then (tag,k) x 3 para.
std::vector<uint8_t> write_dup(const Prep& P, const Residuals& R, const Header& H,
                               const uint8_t* tagk);

}  // namespace stride
