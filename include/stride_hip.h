/*
 * stride_hip.h - C ABI of libstride_hip.so, the MI355X (gfx950) STRIDE codec.
 *
 * Plain C: pointers, sizes and int status codes; no C++ or torch types cross
 * this boundary.  Every function returns STRIDE_OK (0) or a negative status;
 * stride_last_error() copies the message of the last failure on the calling
 * thread.  Device pointers are HIP device memory; `stream` is a hipStream_t
 * (NULL = the default stream).  Handles are not reentrant: one thread at a
 * time per handle; creating handles is thread safe.  The decodes of one
 * handle share its device scratch (the two-pass walk buffer), so they must
 * not overlap on the GPU: issue them on one stream, or order the streams
 * with events.
 *
 * Reference interfaces replaced (file:line in maletsden/meshpress):
 *   decoder  utils/paradelta_v5_cuda.py:580-718  ParaDeltaV5GpuDecoder
 *            bench_cpp/stride_dup_decode_bench.cu:192-509  fused DPR4 kernel
 *            scripts/dump_stride_dup_blob.py:29-152  side-table derivation
 *   encoder  encoder/implementation/stride.py:18-53  STRIDEEncoder.encode
 *            encoder/paradelta_codec.py:758-882  prepare_paradelta_arrays
 *            encoder/paradelta_v5_dup.py:316-464  encode_dup
 *   search   encoder/paradelta_v5_dup.py:159-171  _pick_best_k
 *            encoder/_irlp_fit.py:120-226  _cost_seeds_njit / encoder_cost_batch
 */
#ifndef STRIDE_HIP_H
#define STRIDE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  STRIDE_OK = 0,
  STRIDE_E_ARG = -1,        /* invalid argument            -> ValueError   */
  STRIDE_E_HIP = -2,        /* HIP runtime failure         -> RuntimeError */
  STRIDE_E_MAGIC = -10,     /* bad magic                   -> ValueError   */
  STRIDE_E_VERSION = -11,   /* unsupported version         -> ValueError   */
  STRIDE_E_TRUNCATED = -12, /* stream shorter than fields  -> IndexError   */
  STRIDE_E_CORRUPT = -13,   /* malformed meshlet           -> ValueError   */
  STRIDE_E_UNSUPPORTED = -14,/* valid but outside kernel limits -> ValueError */
  STRIDE_E_PREP = -20       /* encoder front-end failure   -> RuntimeError */
};

/* Copies the calling thread's last error message (NUL terminated, truncated
 * to cap); returns its full length. */
int stride_last_error(char* buf, size_t cap);

/* Selects the device for later calls on this thread (hipSetDevice). */
int stride_init(int device);

/* Library ABI version. */
int stride_version(void);

/* Diagnostic: iteration order of a CPython set of small non-negative ints
 * built by adding in[0..n) in order (the multiseed strip cover keeps the
 * first strictly better seed in that order, utils/amd_gts_nb.py:160-175).
 * Returns the number of distinct keys written to out. */
int stride_debug_pyset_order(const int32_t* in, int32_t n, int32_t* out);

/* ------------------------------------------------------------------------ */
/* Decoder.  Replaces ParaDeltaV5GpuDecoder(data) (utils/paradelta_v5_cuda.py:597)
 * for DPR4 v3 streams as written by encode_dup
 * (encoder/paradelta_v5_dup.py:316-464).  Creation uploads the bytes, runs
 * the index pre-pass kernel (the side table the reference bench reads from a
 * DPRB blob, scripts/dump_stride_dup_blob.py:88-152) and validates every
 * meshlet; decode is then one kernel launch.
 *
 * A stream starting with the PDLT magic (0x504C4454) is decoded as PDLT v5
 * (the format the reference class decodes, utils/paradelta_v5_cuda.py:34-577;
 * CPU decoder encoder/paradelta_v5.py:286-462): creation parses the header
 * and the boundary table on the host and runs the meshlet pre-pass; decode
 * is a boundary-vertex kernel plus one wave-group kernel.  Output equals
 * decode_paradelta_v5 bit for bit (float64 reconstruction): vertices are the
 * boundary table, then each meshlet's interior vertices; indices are global.
 * Limits (STRIDE_E_UNSUPPORTED): n_bnd + n_int <= 256 per meshlet, boundary
 * refs inside the boundary table, residual codes and coder parameters within
 * 32-bit decoding, g_bits <= 32, streams below 512 MiB.  Tune fields used:
 * group_size (1..16) and lds_bytes (4096..65536). */
typedef struct stride_decoder stride_decoder;

int stride_decoder_create(const uint8_t* data, size_t len, stride_decoder** out);
void stride_decoder_destroy(stride_decoder* h);
/* DPRB v4 side-table blob (scripts/dump_stride_dup_blob.py:88-152; the
 * native decode bench's input, bench_cpp/stride_dup_decode_bench.cu:523-595):
 * the DPR4 meshlet body plus per-meshlet tables (bit offsets, n_local, n_tris,
 * n_strips, v_off, t_off, anchor offset, n_kind0, substream offsets).
 * stride_dprb_from_dpr4 derives the tables on the host with the pre-pass
 * parse; with out == NULL only *out_len is set.  stride_decoder_create_dprb
 * decodes a DPRB blob and rejects it (STRIDE_E_CORRUPT) unless every table
 * entry equals what the pre-pass derives from its body. */
int stride_dprb_from_dpr4(const uint8_t* data, size_t len, uint8_t* out, size_t cap,
                          size_t* out_len);
int stride_decoder_create_dprb(const uint8_t* blob, size_t len, stride_decoder** out);
/* SBLB v1 blob (scripts/dump_stride_blob.py:1-114; the reference's native
 * PDLT v5 decode bench input, bench_cpp/stride_decode_bench.cu:430-484): the
 * PDLT v5 stream plus the per-meshlet tables that bench uploads (bit offsets,
 * n_bnd, n_int, n_tris, n_strips, tris_off, int_cursor) and the float32
 * boundary table.  stride_sblb_from_pdlt writes one from a PDLT v5 stream
 * (out == NULL: only *out_len); stride_decoder_create_sblb decodes one and
 * rejects it (STRIDE_E_CORRUPT) unless every table equals what the pre-pass
 * derives from its stream. */
int stride_sblb_from_pdlt(const uint8_t* data, size_t len, uint8_t* out, size_t cap,
                          size_t* out_len);
int stride_decoder_create_sblb(const uint8_t* blob, size_t len, stride_decoder** out);
/* n_v = sum of per-meshlet n_local (duplicated boundary vertices included;
 * PDLT v5: n_boundary + sum of n_int), n_t = triangles, n_meshlets. */
int stride_decoder_sizes(const stride_decoder* h, uint64_t* n_v, uint64_t* n_t,
                         uint32_t* n_meshlets);
/* Decoded world-space positions (float32 xyz, n_v x 3) and uint32 indices
 * (n_t x 3) into caller-owned device buffers; asynchronous on `stream`.  An
 * output may be NULL only when it has no elements (a stream of no meshlets);
 * otherwise STRIDE_E_ARG. */
int stride_decode(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream);
/* Synchronous decode into caller-owned host buffers. */
int stride_decode_to_host(stride_decoder* h, float* verts, uint32_t* tris);
/* The decoder's own device outputs, allocated at creation (float32 n_v x 3,
 * uint32 n_t x 3) and freed by stride_decoder_destroy: a caller that lets
 * the decoder own its outputs, as ParaDeltaV5GpuDecoder.decode() does
 * (utils/paradelta_v5_cuda.py:686-712), passes them to stride_decode (or
 * stride_decode_normalized for PDLT v5). */
int stride_decoder_outputs(stride_decoder* h, float** d_verts, uint32_t** d_tris);
/* stride_decode for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  DPR4 streams only; the decoder's device
 * must be current.  Same kernels and output as stride_decode, but their
 * scratch (the two-pass walk buffer and the work-queue counter) is the
 * library's static 32 MiB device array instead of the decoder's own buffers,
 * and the call allocates nothing.  A decode needing more scratch (64 words
 * + n_v + 4 n_meshlets + 8 in mode 4, 64 otherwise) fails with
 * STRIDE_E_UNSUPPORTED.  Every call on a GPU shares that GPU's one array:
 * calls that overlap on one GPU (other streams, other threads) race. */
int stride_decode_noalloc(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream);
/* PDLT v5 only: normalised positions (float32 of the float64 reconstruction,
 * before scale and center), as the reference's decode() returns them
 * (utils/paradelta_v5_cuda.py:686-712); STRIDE_E_ARG for DPR4. */
int stride_decode_normalized(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream);
/* PDLT v5 reconstruction arithmetic: chain_bits 64 (default) = float64, bit
 * for bit decode_paradelta_v5; 32 = the float32 chain the encoder computed
 * its codes against (encoder/paradelta_v5_nb.py:20-212), so positions are
 * the encoder's own reconstructions and stay within the quantisation bound
 * (the float64 chain drifts from them, by up to 2.5 steps on 16-bit grids).
 * DPR4 accepts only 64. */
int stride_decoder_set_precision(stride_decoder* h, int32_t chain_bits);
/* Stream format: *version = 3 (DPR4 v3) or 5 (PDLT v5); for PDLT v5 also the
 * LIN3 weights and the boundary-table size (zeros for DPR4). */
int stride_decoder_format(const stride_decoder* h, int32_t* version, float* lin3,
                          uint32_t* n_boundary);
/* Bytes resident on the device: compressed stream, derived side table. */
int stride_decoder_bytes(const stride_decoder* h, uint64_t* stream_bytes,
                         uint64_t* side_bytes);
/* Global header fields: center[3], scale, per_coord_err, g_min[3], g_range[3]
 * into f13; g_bits[3], pred_n[9], pred_K[3], code_width into i16. */
int stride_decoder_header(const stride_decoder* h, float* f13, int32_t* i16);

/* Launch configuration (autotuned).
 *   mode 4 (default; tok_chains 16, tok_group 4, group_size 21, lds_bytes
 *     24576): two passes.  First
 *     the token pass, writing the walks to a device buffer and the
 *     triangles to the output: with tok_chains > 0 (the default) each
 *     meshlet's strip tokens are cut into about tok_chains chains, one lane
 *     per chain parses its tokens into 16-bit records and the wave then
 *     forms the triangles and walk entries of its tok_group meshlets; with
 *     tok_chains = 0, one token lane per meshlet and tok_group (<= 64)
 *     consecutive meshlets per wave, with an 8-word seen set per lane.  Then
 *     the axis pass: split groups of up to group_size (<= 21) meshlets,
 *     three lanes per meshlet, staging their residual sections in LDS
 *     (lds_bytes per group) and keeping only the codes there.  The launch
 *     takes the LDS its largest group needs, at most lds_bytes.
 *   mode 2 (group_size 16, lds_bytes 24576 typical): split groups of up to
 *     16 consecutive meshlets per wave, four lanes per meshlet (one decodes
 *     the strip tokens into a walk list in LDS, then one per axis decodes
 *     that axis's residual substreams and prediction chain); lds_bytes of
 *     LDS per group (staged bits + 10 bytes per vertex with uint16 codes,
 *     used when every axis has at most 16 bits, else 16).
 *   mode 1: groups of up to group_size (<= 64) consecutive meshlets per
 *     wave, one lane per meshlet, lds_bytes of LDS per group (staged bits +
 *     packed codes).
 *   In modes 1 and 2, meshlets that do not fit a group, or whose codes leave
 *   the packed range, are decoded by mode 0.
 *   mode 0: one wave per meshlet: waves per workgroup {1,2,4}, LDS stage
 *     words per meshlet {256,512,1024}, use_queue = persistent grid pulling
 *     meshlets from a device atomic counter.
 *   grid: 0 = one workgroup per unit of work, else a persistent grid.
 *   PDLT v5 streams: mode 4 (default: the token pass, then chain groups of
 *   group_size meshlets in lds_bytes, default 16 and 20480) or mode 3 (one
 *   kernel).  The token
 *   pass runs token chains as for DPR4 (default tok_chains 16, tok_group 4)
 *   or, with tok_chains = 0, one token lane per meshlet (stage_words = its
 *   LDS words, tok_group <= 64 meshlets per wave).  A stream with a meshlet
 *   too large for a chain token wave (48 KiB of LDS with its 16-bit token
 *   records) runs the token lanes, 64 meshlets per wave, and get_tune
 *   reports tok_chains 0 and tok_group 64.
 *   get_tune reports in `reserved` the meshlets decoded by mode 0. */
typedef struct stride_tune {
  int32_t waves;
  int32_t stage_words;
  int32_t grid;
  int32_t use_queue;
  int32_t mode;
  int32_t group_size;
  int32_t lds_bytes;
  int32_t reserved;
  int32_t tok_group;   /* mode 4: meshlets per token-pass wave, 1..64 (0: the
                          default: 4 for chain token waves, 32 for DPR4 token
                          lanes, 64 for PDLT v5 token lanes) */
  int32_t tok_chains;  /* mode 4: 0 = one token lane per meshlet; 1..16 =
                          the chain token pass, each meshlet's strip tokens cut
                          into about tok_chains chains of >= 16 tokens parsed
                          by one lane each (up to 64 chains and tok_group
                          meshlets per wave), then triangles and walk entries
                          formed by the whole wave per meshlet.  -1 = keep. */
} stride_tune;

int stride_decoder_set_tune(stride_decoder* h, const stride_tune* t);
int stride_decoder_get_tune(const stride_decoder* h, stride_tune* t);

/* ------------------------------------------------------------------------ */
/* Batched decode: many DPR4 streams in one device buffer, decoded by one
 * launch over a global meshlet table (no reference counterpart beyond the
 * per-mesh loop of utils/paradelta_v5_cuda.py:686-712).  Output is the
 * concatenation of every mesh's decode_dup result; indices stay relative
 * to each mesh's own vertex array. */
typedef struct stride_batch stride_batch;

int stride_batch_create(const uint8_t* const* blobs, const size_t* lens, int32_t n_meshes,
                        stride_batch** out);
/* stride_batch_create, then stride_batch_set_tune(tune) before any decode
 * (tune == NULL: the defaults). */
int stride_batch_create_tuned(const uint8_t* const* blobs, const size_t* lens, int32_t n_meshes,
                              const stride_tune* tune, stride_batch** out);
void stride_batch_destroy(stride_batch* h);
/* The batch's own device outputs (every mesh's vertices, then triangles, in
 * mesh order; stride_batch_offsets), as stride_decoder_outputs. */
int stride_batch_outputs(stride_batch* h, float** d_verts, uint32_t** d_tris);
int stride_batch_sizes(const stride_batch* h, uint64_t* n_v, uint64_t* n_t,
                       uint32_t* n_meshlets);
/* Per mesh first vertex / first triangle in the concatenated outputs,
 * n_meshes + 1 entries each. */
int stride_batch_offsets(const stride_batch* h, uint64_t* v_base, uint64_t* t_base);
int stride_batch_bytes(const stride_batch* h, uint64_t* stream_bytes, uint64_t* side_bytes);
int stride_batch_set_tune(stride_batch* h, const stride_tune* t);
int stride_batch_get_tune(const stride_batch* h, stride_tune* t);
/* One launch of each decode pass over all meshlets of all meshes. */
int stride_batch_decode(stride_batch* h, float* d_verts, uint32_t* d_tris, void* stream);
/* Same launch replayed from a captured hipGraph (captured on first use for
 * this output pair and tune). */
int stride_batch_decode_graph(stride_batch* h, float* d_verts, uint32_t* d_tris, void* stream);
/* Baseline: one launch per mesh, in order. */
int stride_batch_decode_per_mesh(stride_batch* h, float* d_verts, uint32_t* d_tris,
                                 void* stream);

/* ------------------------------------------------------------------------ */
/* Encoder front-end (CPU).  Restates prepare_paradelta_arrays
 * (encoder/paradelta_codec.py:758-882) from the partition on: global
 * triangle adjacency, joint-learned meshlet growth
 * (utils/meshlet_gen_joint_nb.py:66-430), boundary Morton order, per-meshlet
 * plan (encoder/paradelta_codec.py:347-439), multiseed strips
 * (utils/amd_gts_nb.py:150-175), and the walk (encoder/paradelta_v5_dup.py:114-152).
 * Inputs: tris n_t x 3, normalised positions vn n_v x 3, face normals
 * fn n_t x 3, global codes n_v x 3, cost constants k6/nm6, strip seed table
 * (6 ints per row, row = triangle count). */
typedef struct stride_encoder stride_encoder;

int stride_encoder_create(const int64_t* tris, int64_t n_t, const double* vn, int64_t n_v,
                          const double* fn, const int64_t* codes, int32_t max_verts,
                          int32_t max_tris, const double* k6, const double* nm6,
                          const int32_t* seed_table, int32_t seed_rows, int32_t n_threads,
                          stride_encoder** out);
void stride_encoder_destroy(stride_encoder* h);
int stride_encoder_counts(const stride_encoder* h, int64_t* n_meshlets, int64_t* n_local_total,
                          int64_t* n_tris_total, int64_t* n_strips_total,
                          int64_t* n_para_total);
/* Plans as flat arrays (meshlet-major); walk5 = (v, kind, a, b, c) bytes
 * per local vertex in walk order. */
int stride_encoder_export_plans(const stride_encoder* h, int32_t* n_local, int32_t* n_tris,
                                int32_t* n_strips, int32_t* n_bnd, int64_t* l2g,
                                uint8_t* tris_local, int32_t* strip_len, int32_t* strip_tris,
                                uint8_t* walk5);
/* IRLP fit samples (encoder/_irlp_fit.py:72-97): global codes of (a, b, c, v)
 * for every parallelogram vertex, n_para x 3 each, and per-meshlet offsets
 * (n_meshlets + 1). */
int stride_encoder_para_samples(const stride_encoder* h, int32_t* A, int32_t* B, int32_t* C,
                                int32_t* V, int64_t* off);
/* Residuals for predictor (pred_n 3x3, pred_K 3) (paradelta_v5_dup.py:343-392).
 * Returns the number of residual values and max |residual|. */
int stride_encoder_residuals(stride_encoder* h, const int32_t* pred_n, const int32_t* pred_K,
                             int64_t* n_values, int64_t* max_abs);
/* Residual streams: 6 per meshlet (dx, dy, dz, px, py, pz), int32 values and
 * 6 * n_meshlets + 1 offsets. */
int stride_encoder_export_residuals(const stride_encoder* h, int32_t* vals,
                                    int64_t* stream_off);
/* Per-meshlet features of the learned proposal, 32 float32 per meshlet of
 * sel[0..n_sel) (every meshlet when sel is NULL) (no
 * reference counterpart: they feed the MLP that replaces the search of
 * encoder/_irlp_fit.py:364-394 and paradelta_v5_dup.py:159-171).  Stream
 * statistics come from the canonical predictor a + b - c (source 0) or from
 * the residuals of the last stride_encoder_residuals call (source 1);
 * streams s = dx dy dz px py pz, u = zigzag residual:
 *   0 n_local/256   1 n_tris/256   2 n_strips/n_tris   3 n_kind0/n_local
 *   4..6   log2(1 + local code extent of axis d) / 16
 *   7..10  fraction of triangles in strips of length 1, 2-3, 4-15, >= 16
 *   11..16 log2(1 + count_s) / 8
 *   17..22 mean of floor(log2(u + 1)) over stream s, / 16
 *   23..28 log2(1 + mean u over stream s) / 16
 *   29..31 clamp(sum(x*y) / sum(x*x), 0, 2) / 2 over the parallelogram
 *          vertices of axis d, x = a + b - 2c, y = v - c (1 when sum(x*x) = 0) */
int stride_encoder_features(const stride_encoder* h, int32_t source, const int32_t* sel,
                            int64_t n_sel, float* out);
/* The IRLP fit samples of stride_encoder_para_samples for meshlets
 * sel[0..n_sel) (all meshlets when sel is NULL), as axis-major int32
 * columns: cols[(q * 3 + d) * n + i] for q = a, b, c, v, axis d, sample i,
 * n = off[rows]; off gets rows + 1 offsets.  With cols NULL only off is
 * written (to size cols).  The subsample of encoder/_irlp_fit.py:72-97 is a
 * selection of meshlets. */
int stride_encoder_para_columns(const stride_encoder* h, const int32_t* sel, int64_t n_sel,
                                int32_t* cols, int64_t* off);
/* DPR4 v3 bytes (paradelta_v5_dup.py:239-313,412-455).  f13 = center[3],
 * scale, per_coord_err, g_min[3], g_range[3] (float32 values written
 * verbatim); g_bits[3]; tagk = 12 bytes per meshlet ((tag, k) x 3 delta then
 * x 3 para).  With out == NULL only *len is set. */
int stride_encoder_write(const stride_encoder* h, const float* f13, const int32_t* g_bits,
                         const int32_t* pred_n, const int32_t* pred_K, const uint8_t* tagk,
                         uint8_t* out, size_t cap, size_t* len);

/* PDLT v5 writer over the same plans (encode_from_prepared_v5,
 * encoder/paradelta_v5.py:144-283, with the float32 chain and meshlet
 * layout of encoder/paradelta_v5_nb.py:20-539).  Three calls:
 *   _v5_prepare: vn = normalised positions (n_v x 3 f64, as passed to
 *     stride_encoder_create), f13 / g_bits as for stride_encoder_write.
 *     Runs the plain-parallelogram pass; returns the boundary-table size and
 *     the number of LIN3 fit samples;
 *   _v5_samples: the samples as four n_samples x 3 f64 arrays (a, b, c
 *     reconstructions and the true position), for the least-squares fit
 *     (_fit_linear3, encoder/paradelta_codec.py:626-639) done by the caller;
 *   _write_v5: the LIN3 pass with the fitted float32 weights and the bytes
 *     (out == NULL: only *len). */
int stride_encoder_v5_prepare(stride_encoder* h, const double* vn, const float* f13,
                              const int32_t* g_bits, int64_t* n_boundary, int64_t* n_samples);
int stride_encoder_v5_samples(const stride_encoder* h, double* A, double* B, double* C,
                              double* T);
int stride_encoder_write_v5(const stride_encoder* h, const float* lin3, uint8_t* out, size_t cap,
                            size_t* len);

/* ------------------------------------------------------------------------ */
/* GPU search kernels (device pointers). */

/* Exhaustive per-stream coder choice (_pick_best_k, paradelta_v5_dup.py:159-171):
 * Rice k 0..11 and Exp-Golomb k 0..7 over the zigzag residuals of each of
 * n_streams streams [off[s], off[s+1]); lowest k wins ties, Rice wins ties
 * against Exp-Golomb, empty streams -> (1, 0).  Writes tag,k bytes
 * (2 per stream) and the winning body bit count. */
int stride_select_codes(const int32_t* d_vals, const int64_t* d_off, int64_t n_streams,
                        uint8_t* d_tagk, int64_t* d_bits, void* stream);
/* All 20 candidate body costs per stream (n_streams x 20 int64: Rice k
 * 0..11 then Exp-Golomb k 0..7); the exhaustive labeller. */
int stride_stream_costs(const int32_t* d_vals, const int64_t* d_off, int64_t n_streams,
                        int64_t* d_costs, void* stream);
/* IRLP candidate cost (_cost_seeds_njit, encoder/_irlp_fit.py:120-183): for
 * each of S candidates (n0, n1, n2) with shift K, the sum over meshlets of
 * min over the 20 coders of the parallelogram-residual bits of one axis.
 * a, b, c, v: n samples; off: n_m + 1 meshlet offsets. */
int stride_irlp_cost(const int32_t* d_a, const int32_t* d_b, const int32_t* d_c,
                     const int32_t* d_v, const int64_t* d_off, int64_t n_m,
                     const int32_t* d_cand, int32_t S, int32_t K, int64_t* d_cost,
                     void* stream);
/* The same costs kept per meshlet: d_cost_m[m * S + s] (n_m x S int64), the
 * labeller of the learned predictor proposal (no reference counterpart;
 * the terms are those summed by _cost_seeds_njit). */
int stride_irlp_cost_meshlets(const int32_t* d_a, const int32_t* d_b, const int32_t* d_c,
                              const int32_t* d_v, const int64_t* d_off, int64_t n_m,
                              const int32_t* d_cand, int32_t S, int32_t K, int64_t* d_cost_m,
                              void* stream);

/* ------------------------------------------------------------------------ */
/* Learned proposal network (device pointers; bf16 buffers are uint16_t bit
 * patterns).  It replaces the exhaustive searches of
 * encoder/_irlp_fit.py:364-394 and paradelta_v5_dup.py:159-171 by a proposal
 * that the search kernels above then score; the reference has no network.
 * MLP 32 -> 128 -> 128 -> 176, ReLU, BF16 operands, FP32 accumulation on
 * the matrix cores.  Parameters (float32, n_params): W1 32x128 | W2 128x128 |
 * W3 128x176 | b1 | b2 | b3; d_wb holds bf16 copies of the n_weights
 * weights.  Logit heads: predictor 3 x 17 (columns 0..50), coders 6 x 20
 * (51..170); 171..175 are padding. */
int stride_mlp_sizes(int32_t* n_in, int32_t* n_hidden, int32_t* n_out, int32_t* n_weights,
                     int32_t* n_params);
/* C (M x N) = op(A) op(B) with op(A)[m][k] = trans_a ? A[k*lda+m] : A[m*lda+k]
 * (likewise B).  epi 0: float C (+ bias[n] if bias); 1: bf16 relu(C + bias);
 * 2: bf16 C where mask[m*ldm+n] > 0, else 0. */
int stride_mlp_gemm(const uint16_t* d_a, int32_t lda, int32_t trans_a, const uint16_t* d_b,
                    int32_t ldb, int32_t trans_b, int32_t M, int32_t N, int32_t K, int32_t epi,
                    const float* d_bias, const uint16_t* d_mask, int32_t ldm, float* d_cf,
                    uint16_t* d_cb, int32_t ldc, void* stream);
/* x (n x 32 bf16) -> h1, h2 (n x 128 bf16), z (n x 176 float logits). */
int stride_mlp_forward(const uint16_t* d_x, int64_t n, const uint16_t* d_wb, const float* d_params,
                       uint16_t* d_h1, uint16_t* d_h2, float* d_z, void* stream);
/* Grouped softmax cross-entropy over the 9 heads; labels n x 9 (255 = head
 * ignored).  dz = scale * (softmax - onehot) as bf16; loss per sample. */
int stride_mlp_xent(const float* d_z, const uint8_t* d_labels, int64_t n, float scale,
                    uint16_t* d_dz, float* d_loss, void* stream);
/* Gradient of the loss behind dz into d_grad (n_params floats); d_dh2 and
 * d_dh1 (n x 128 bf16) receive the hidden-layer gradients. */
int stride_mlp_backward(const uint16_t* d_x, int64_t n, const uint16_t* d_wb, const uint16_t* d_h1,
                        const uint16_t* d_h2, const uint16_t* d_dz, uint16_t* d_dh2,
                        uint16_t* d_dh1, float* d_grad, void* stream);
/* Adam step `step` (>= 1) on the float32 parameters; refreshes d_wb. */
int stride_mlp_adam(float* d_params, float* d_m, float* d_v, const float* d_grad, uint16_t* d_wb,
                    int64_t n_params, float lr, float beta1, float beta2, float eps, int32_t step,
                    void* stream);
/* float -> bf16, round to nearest even. */
int stride_mlp_to_bf16(const float* d_x, int64_t n, uint16_t* d_y, void* stream);
/* Inference: predictor-head probabilities (n x 51 float), coder-head arg-max
 * class (n x 6) and its probability (n x 6). */
int stride_mlp_heads(const float* d_z, int64_t n, float* d_pred_prob, uint8_t* d_cls,
                     float* d_cls_prob, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* STRIDE_HIP_H */
