// Native decode bench over the C ABI (the counterpart of the reference's
// bench_cpp/stride_dup_decode_bench.cu and stride_decode_bench.cu command
// lines): loads a DPR4 v3 stream, a DPRB v4 blob, a PDLT v5 stream or an
// SBLB v1 blob, decodes it repeatedly into device buffers and prints one
// JSON line with the event-timed decode rate.
//
//   stride_bench FILE [--iters N] [--warmup W] [--mode M] [--group G]
//                [--lds BYTES] [--grid R] [--queue 0|1] [--stage WORDS] [--tok N]
//                [--chains K] [--chain 64|32]
//
// --tok sets the meshlets per token-pass wave of mode 4, --chains the token
// chains a meshlet is cut into (0: one token lane per meshlet).  PDLT v5
// uses --mode (3 one kernel, 4 token pass + chain pass), --group, --lds,
// --stage (the token lanes' LDS words), --tok, --chains and --chain
// (float64 or float32 chain).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stride_hip.h"

static int die(const char* what, int rc) {
  char msg[512];
  stride_last_error(msg, sizeof msg);
  std::fprintf(stderr, "stride_bench: %s failed (%d): %s\n", what, rc, msg);
  return 1;
}

#define HIP_OK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "stride_bench: %s: %s\n", #x, hipGetErrorString(e_));     \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s FILE [--iters N] [--warmup W] [--mode M] [--group G] "
                         "[--lds BYTES] [--grid R] [--queue 0|1] [--tok T] [--chains K] [--stage W] "
                         "[--chain 64|32]\n", argv[0]);
    return 2;
  }
  int iters = 100, warmup = 10, chain = 64;
  stride_tune want{};
  bool set_mode = false, set_group = false, set_lds = false, set_grid = false, set_queue = false,
       set_stage = false, set_tok = false, set_chains = false;
  for (int i = 2; i + 1 < argc; i += 2) {
    const std::string k = argv[i];
    const int v = std::atoi(argv[i + 1]);
    if (k == "--iters") iters = v;
    else if (k == "--warmup") warmup = v;
    else if (k == "--mode") { want.mode = v; set_mode = true; }
    else if (k == "--group") { want.group_size = v; set_group = true; }
    else if (k == "--lds") { want.lds_bytes = v; set_lds = true; }
    else if (k == "--grid") { want.grid = v; set_grid = true; }
    else if (k == "--queue") { want.use_queue = v; set_queue = true; }
    else if (k == "--chain") chain = v;
    else if (k == "--stage") { want.stage_words = v; set_stage = true; }
    else if (k == "--tok") { want.tok_group = v; set_tok = true; }
    else if (k == "--chains") { want.tok_chains = v; set_chains = true; }
    else { std::fprintf(stderr, "stride_bench: unknown option %s\n", k.c_str()); return 2; }
  }
  if (iters < 1 || warmup < 0) { std::fprintf(stderr, "stride_bench: bad iteration counts\n"); return 2; }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) { std::perror(argv[1]); return 1; }
  std::vector<uint8_t> data;
  uint8_t buf[1 << 16];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  const bool dprb = data.size() >= 4 && std::memcmp(data.data(), "DPRB", 4) == 0;
  const bool sblb = data.size() >= 4 && std::memcmp(data.data(), "SBLB", 4) == 0;
  int rc = stride_init(0);
  if (rc) return die("stride_init", rc);
  stride_decoder* h = nullptr;
  rc = dprb ? stride_decoder_create_dprb(data.data(), data.size(), &h)
            : sblb ? stride_decoder_create_sblb(data.data(), data.size(), &h)
                   : stride_decoder_create(data.data(), data.size(), &h);
  if (rc) return die("decoder create", rc);
  int32_t version = 3;
  stride_decoder_format(h, &version, nullptr, nullptr);
  if ((rc = stride_decoder_set_precision(h, chain))) return die("set_precision", rc);
  const char* fmt = dprb ? "dprb" : sblb ? "sblb" : version == 5 ? "pdlt_v5" : "dpr4";
  stride_tune t{};
  stride_decoder_get_tune(h, &t);
  if (set_mode) t.mode = want.mode;
  if (set_group) t.group_size = want.group_size;
  if (set_lds) t.lds_bytes = want.lds_bytes;
  if (set_grid) t.grid = want.grid;
  if (set_queue) t.use_queue = want.use_queue;
  if (set_stage) t.stage_words = want.stage_words;
  if (set_tok) t.tok_group = want.tok_group;
  if (set_chains) t.tok_chains = want.tok_chains;
  if ((rc = stride_decoder_set_tune(h, &t))) return die("set_tune", rc);
  uint64_t n_v = 0, n_t = 0, sb = 0, side = 0;
  uint32_t n_m = 0;
  stride_decoder_sizes(h, &n_v, &n_t, &n_m);
  stride_decoder_bytes(h, &sb, &side);
  float* d_v = nullptr;
  uint32_t* d_t = nullptr;
  HIP_OK(hipMalloc(&d_v, 12 * (n_v ? n_v : 1)));
  HIP_OK(hipMalloc(&d_t, 12 * (n_t ? n_t : 1)));
  hipStream_t s;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  for (int i = 0; i < warmup; ++i)
    if ((rc = stride_decode(h, d_v, d_t, s))) return die("decode", rc);
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, s));
  for (int i = 0; i < iters; ++i)
    if ((rc = stride_decode(h, d_v, d_t, s))) return die("decode", rc);
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  float ms_total = 0.f;
  HIP_OK(hipEventElapsedTime(&ms_total, e0, e1));
  const double ms = ms_total / iters;
  const double bytes = (double)sb + (double)side + 12.0 * (double)n_v + 12.0 * (double)n_t;
  stride_decoder_get_tune(h, &t);
  std::printf("{\"file\": \"%s\", \"format\": \"%s\", \"n_tris\": %llu, \"n_verts\": %llu, "
              "\"meshlets\": %u, \"iters\": %d, \"ms\": %.6f, \"gtris_s\": %.4f, \"hbm_gb_s\": %.2f, "
              "\"chain\": %d, \"tune\": {\"mode\": %d, \"group_size\": %d, \"lds_bytes\": %d, "
              "\"grid\": %d, \"use_queue\": %d, \"stage_words\": %d, \"tok_group\": %d, "
              "\"tok_chains\": %d}}\n",
              argv[1], fmt, (unsigned long long)n_t, (unsigned long long)n_v,
              n_m, iters, ms, (double)n_t / (ms * 1e-3) / 1e9, bytes / (ms * 1e-3) / 1e9, chain, t.mode,
              t.group_size, t.lds_bytes, t.grid, t.use_queue, t.stage_words, t.tok_group,
              t.tok_chains);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(d_v);
  (void)hipFree(d_t);
  stride_decoder_destroy(h);
  return 0;
}
