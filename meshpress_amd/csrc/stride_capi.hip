// C ABI of libstride_hip.so: decoder / batch handles, encoder front-end
// handle.  Host code; kernels live in stride_decode.hip / stride_search.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/stride_hip.h"
#include "pdlt_decode.h"
#include "stride_decode.h"
#include "stride_device.h"
#include "stride_err.h"
#include "stride_host.h"
#include "stride_lane.h"

using namespace stride;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

int stride::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return fail(STRIDE_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e));    \
  } while (0)

uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
float bef32(const uint8_t* p) {
  uint32_t u = be32(p);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// DPR4 v3 global header (encoder/paradelta_v5_dup.py:412-446), 86 bytes,
// every field byte aligned and big endian.
struct DupHeader {
  uint32_t magic, version, code_width;
  float center[3], scale, pce, g_min[3], g_range[3];
  int32_t g_bits[3], pred_n[9], pred_K[3];
  uint32_t n_v, n_t, n_m;
};

int parse_header(const uint8_t* d, size_t len, DupHeader& h) {
  if (len < 5) return fail(STRIDE_E_TRUNCATED, "stream shorter than the DPR4 header");
  h.magic = be32(d);
  if (h.magic != kMagicDup) {
    char b[64];
    std::snprintf(b, sizeof b, "bad magic 0x%08X", h.magic);
    return fail(STRIDE_E_MAGIC, b);
  }
  h.version = d[4];
  if (h.version != kVersionDup) {
    char b[64];
    std::snprintf(b, sizeof b, "expected v%u, got v%u", kVersionDup, h.version);
    return fail(STRIDE_E_VERSION, b);
  }
  if (len < 86) return fail(STRIDE_E_TRUNCATED, "stream shorter than the DPR4 header");
  h.code_width = d[5];
  for (int i = 0; i < 3; ++i) h.center[i] = bef32(d + 6 + 4 * i);
  h.scale = bef32(d + 18);
  h.pce = bef32(d + 22);
  for (int i = 0; i < 3; ++i) h.g_min[i] = bef32(d + 26 + 4 * i);
  for (int i = 0; i < 3; ++i) h.g_range[i] = bef32(d + 38 + 4 * i);
  for (int i = 0; i < 3; ++i) h.g_bits[i] = d[50 + i];
  for (int i = 0; i < 9; ++i) {
    uint32_t u = ((uint32_t)d[53 + 2 * i] << 8) | d[54 + 2 * i];
    h.pred_n[i] = (u & 0x8000u) ? (int32_t)u - 0x10000 : (int32_t)u;
  }
  for (int i = 0; i < 3; ++i) h.pred_K[i] = d[71 + i];
  h.n_v = be32(d + 74);
  h.n_t = be32(d + 78);
  h.n_m = be32(d + 82);
  if ((uint64_t)len < 86ull + 4ull * h.n_m)
    return fail(STRIDE_E_TRUNCATED, "stream shorter than its meshlet offset table");
  for (int i = 0; i < 3; ++i)
    if (h.g_bits[i] < 1 || h.g_bits[i] > 30)
      return fail(STRIDE_E_UNSUPPORTED, "g_bits outside [1, 30] is not supported by the GPU decoder");
  for (int i = 0; i < 3; ++i)
    if (h.pred_K[i] > 30)
      return fail(STRIDE_E_UNSUPPORTED, "predictor shift K > 30 is not supported");
  return STRIDE_OK;
}

const char* scan_err_text(uint32_t e) {
  switch (e) {
    case kScanErrSize: return "meshlet counts outside the supported range (n_local, n_tris <= 256)";
    case kScanErrOverrun: return "meshlet fields run past the meshlet's bit range";
    case kScanErrFifo: return "reuse FIFO index past the FIFO fill";
    case kScanErrIndex: return "local vertex id >= n_local";
    case kScanErrStrips: return "strip lengths do not add up to n_tris";
    case kScanErrWalk: return "strip walk does not cover every local vertex";
    case kScanErrTag: return "residual coder parameter k > 24 is not supported";
    default: return "malformed meshlet";
  }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Scratch of stride_decode_noalloc: the work-queue counter (words 0..63)
// and the two-pass walk slots, 32 MiB in all.  Each GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not the device array: kernels and copies get the
// address hipGetSymbolAddress returns for the current device (scratch_ptr).
constexpr uint64_t kScratchWords = 8ull << 20;
__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current
// device; looked up once per device.
int scratch_ptr(int device, uint32_t** out) {
  if (device < 0 || device >= kMaxDevices) return fail(STRIDE_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(STRIDE_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_dev[device];
  return STRIDE_OK;
}

}  // namespace

// One device-resident set of DPR4 streams with its side table.
struct DecodeSet {
  int device = 0;
  std::vector<DupHeader> hdr;
  std::vector<uint64_t> v_base, t_base;   // n_meshes + 1
  std::vector<uint32_t> m_base;           // first meshlet of each mesh, n_meshes + 1
  std::vector<MeshletEntry> tab;          // host copy of the side table
  std::vector<MeshParams> mp;
  uint8_t* d_buf = nullptr;
  size_t buf_bytes = 0, stream_bytes = 0;
  MeshletEntry* d_tab = nullptr;
  MeshParams* d_mesh = nullptr;
  uint32_t* d_queue = nullptr;
  uint32_t n_m = 0;
  uint64_t n_v = 0, n_t = 0;
  float* d_v_own = nullptr;
  uint32_t* d_t_own = nullptr;
  DecodeConfig cfg;
  // group layout (mode 1)
  GroupEntry* d_groups = nullptr;
  uint32_t n_groups = 0;
  std::vector<uint32_t> g_base;           // first group of each mesh, n_meshes + 1
  uint32_t* d_legacy = nullptr;           // meshlets decoded by the wave-per-meshlet kernel
  // mode 4: token groups (up to 64 meshlets, unions of consecutive split
  // groups) and the walks, one word per local vertex
  GroupEntry* d_tgroups = nullptr;
  uint32_t n_tgroups = 0, tlds_bytes = 0;
  uint32_t group_lds = 0;                 // LDS bytes of the group launches (largest group)
  std::vector<uint32_t> tg_base;          // first token group of each mesh, n_meshes + 1
  uint32_t* d_walk = nullptr;
  ChainEntry* d_chains = nullptr;         // token chains (tok_chains > 0)
  uint32_t n_chains = 0;
  uint32_t n_legacy = 0;
  std::vector<uint32_t> l_base;           // first legacy entry of each mesh
  bool ready = false;                     // the launch layout is complete on the device
  bool narrow_codes = true;               // narrow(), fixed by layout()
  // captured graphs keyed by (verts, tris, tune)
  std::map<std::tuple<void*, void*, int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
  hipStream_t cap_stream = nullptr;

  ~DecodeSet() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    (void)hipFree(d_buf);
    (void)hipFree(d_tab);
    (void)hipFree(d_mesh);
    (void)hipFree(d_queue);
    (void)hipFree(d_v_own);
    (void)hipFree(d_t_own);
    (void)hipFree(d_groups);
    (void)hipFree(d_legacy);
    (void)hipFree(d_tgroups);
    (void)hipFree(d_walk);
    (void)hipFree(d_chains);
  }

  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  int build(const uint8_t* const* blobs, const size_t* lens, int n_meshes) {
    if (n_meshes < 1) return fail(STRIDE_E_ARG, "need at least one stream");
    hdr.resize(n_meshes);
    std::vector<uint64_t> blob_byte(n_meshes);
    uint64_t total = 0;
    uint64_t nm_total = 0;
    for (int i = 0; i < n_meshes; ++i) {
      if (!blobs[i] && lens[i]) return fail(STRIDE_E_ARG, "null stream pointer");
      int rc = parse_header(blobs[i], lens[i], hdr[i]);
      if (rc) return rc;
      blob_byte[i] = total;
      total += (lens[i] + 15) & ~(size_t)15;
      nm_total += hdr[i].n_m;
      stream_bytes += lens[i];
    }
    if (nm_total > 0x7FFFFFFFull) return fail(STRIDE_E_UNSUPPORTED, "too many meshlets");
    if (total > 0x3FFFFFFF0ull) return fail(STRIDE_E_UNSUPPORTED, "streams larger than 16 GiB");
    n_m = (uint32_t)nm_total;
    buf_bytes = total + 64;
    std::vector<uint8_t> hbuf(buf_bytes, 0);
    tab.assign(n_m, MeshletEntry{});
    mp.assign(n_meshes, MeshParams{});
    m_base.assign(n_meshes + 1, 0);
    uint32_t m = 0;
    for (int i = 0; i < n_meshes; ++i) {
      std::memcpy(hbuf.data() + blob_byte[i], blobs[i], lens[i]);
      const DupHeader& h = hdr[i];
      MeshParams& p = mp[i];
      std::memset(&p, 0, sizeof p);
      p.blob_byte = blob_byte[i];
      for (int d = 0; d < 3; ++d) {
        p.g_min[d] = h.g_min[d];
        p.g_range[d] = h.g_range[d];
        p.center[d] = h.center[d];
        p.g_bits[d] = h.g_bits[d];
        p.pred_K[d] = h.pred_K[d];
      }
      for (int j = 0; j < 9; ++j) p.pred_n[j] = h.pred_n[j];
      p.scale = h.scale;
      m_base[i] = m;
      const uint64_t total_bits = 8ull * lens[i];
      const uint64_t region = 8ull * (86ull + 4ull * h.n_m);
      if (total_bits > 0xFFFFFFF0ull) return fail(STRIDE_E_UNSUPPORTED, "stream larger than 512 MiB");
      for (uint32_t j = 0; j < h.n_m; ++j) {
        uint32_t off = be32(blobs[i] + 86 + 4 * j);
        uint64_t nxt = j + 1 < h.n_m ? be32(blobs[i] + 86 + 4 * (j + 1)) : total_bits;
        if ((j == 0 && off != region) || off >= nxt || nxt > total_bits)
          return fail(STRIDE_E_CORRUPT, "meshlet offset table is not consistent with the stream");
        MeshletEntry& e = tab[m + j];
        e.mesh = (uint32_t)i;
        e.start_bit = off;
        e.end_bit = (uint32_t)nxt;
      }
      m += h.n_m;
    }
    m_base[n_meshes] = m;
    // Host-side validation is complete; from here on the device is used.
    HIP_TRY(hipGetDevice(&device));
    HIP_TRY(hipMalloc(&d_buf, buf_bytes));
    HIP_TRY(hipMemcpy(d_buf, hbuf.data(), buf_bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&d_tab, sizeof(MeshletEntry) * std::max<uint32_t>(n_m, 1)));
    HIP_TRY(hipMalloc(&d_mesh, sizeof(MeshParams) * n_meshes));
    HIP_TRY(hipMalloc(&d_queue, 256));
    HIP_TRY(hipMemcpy(d_mesh, mp.data(), sizeof(MeshParams) * n_meshes, hipMemcpyHostToDevice));
    std::vector<MeshletScan> scan(n_m);
    if (n_m) {
      MeshletScan* d_scan = nullptr;
      HIP_TRY(hipMemcpy(d_tab, tab.data(), sizeof(MeshletEntry) * n_m, hipMemcpyHostToDevice));
      HIP_TRY(hipMalloc(&d_scan, sizeof(MeshletScan) * n_m));
      hipError_t e1 = launch_scan(d_buf, d_tab, d_mesh, n_m, d_scan, nullptr);
      hipError_t e2 = e1 == hipSuccess ? hipDeviceSynchronize() : e1;
      hipError_t e3 = e2 == hipSuccess
                          ? hipMemcpy(scan.data(), d_scan, sizeof(MeshletScan) * n_m,
                                      hipMemcpyDeviceToHost)
                          : e2;
      (void)hipFree(d_scan);
      HIP_TRY(e3);
    }
    // Side table: prefix sums, validation of meshlet contiguity.
    v_base.assign(n_meshes + 1, 0);
    t_base.assign(n_meshes + 1, 0);
    uint64_t vt = 0, tt = 0;
    for (int i = 0; i < n_meshes; ++i) {
      uint64_t vo = 0, to = 0;
      v_base[i] = vt;
      t_base[i] = tt;
      mp[i].v_base = vt;
      mp[i].t_base = tt;
      for (uint32_t j = m_base[i]; j < m_base[i + 1]; ++j) {
        const MeshletScan& s = scan[j];
        if (s.error) {
          char b[160];
          std::snprintf(b, sizeof b, "meshlet %u of stream %d: %s", j - m_base[i], i,
                        scan_err_text(s.error));
          return fail(STRIDE_E_CORRUPT, b);
        }
        MeshletEntry& e = tab[j];
        if (j + 1 < m_base[i + 1] && s.end_bit != e.end_bit)
          return fail(STRIDE_E_CORRUPT,
                      "meshlet sizes disagree with the meshlet offset table");
        if (vo > 0xFFFFFFFFull - 256 || to > 0xFFFFFFFFull - 256)
          return fail(STRIDE_E_UNSUPPORTED, "mesh too large for 32-bit indices");
        e.v_off = (uint32_t)vo;
        e.t_off = (uint32_t)to;
        e.n_local = s.n_local;
        e.n_tris = s.n_tris;
        e.n_strips = s.n_strips | (s.flags & kFlagRawInFifo);
        e.n_kind0 = s.n_kind0;
        e.resid_bit = s.resid_bit;
        for (int k = 0; k < 6; ++k) e.sub_bit[k] = s.sub_bit[k];
        e.end_bit = s.end_bit;
        vo += s.n_local;
        to += s.n_tris;
      }
      vt += vo;
      tt += to;
    }
    v_base[n_meshes] = vt;
    t_base[n_meshes] = tt;
    n_v = vt;
    n_t = tt;
    HIP_TRY(hipMemcpy(d_mesh, mp.data(), sizeof(MeshParams) * n_meshes, hipMemcpyHostToDevice));
    if (n_m)
      HIP_TRY(hipMemcpy(d_tab, tab.data(), sizeof(MeshletEntry) * n_m, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&d_v_own, 12 * std::max<uint64_t>(n_v, 1)));
    HIP_TRY(hipMalloc(&d_t_own, 12 * std::max<uint64_t>(n_t, 1)));
    // default launch shape (the fastest measured on the 1.3M- and
    // 7.2M-triangle meshes and the 2,048-mesh batch): 16 token chains per
    // meshlet, 4 meshlets per token wave, axis groups of 21 meshlets
    cfg.tok_chains = 16;
    cfg.tok_group = 4;
    cfg.group_size = kAxisSlots;
    cfg.lds_bytes = 24576;
    return layout();
  }

  // Partition the meshlets into groups for the current tune and run one
  // validation decode: groups whose codes leave the packed range move to
  // the wave-per-meshlet list, which must itself fit 32-bit codes.
  int layout() {
    // not ready until every table of the new layout is on the device: a
    // failure on the way leaves the set refusing to launch
    ready = false;
    narrow_codes = narrow();
    drop_graphs();
    (void)hipFree(d_groups);
    (void)hipFree(d_legacy);
    d_groups = nullptr;
    d_legacy = nullptr;
    const int n_meshes = (int)mp.size();
    std::vector<uint8_t> legacy(n_m, cfg.mode == 0 ? 1 : 0);
    std::vector<GroupEntry> groups;
    for (int pass = 0; pass < 2; ++pass) {
      groups.clear();
      g_base.assign(n_meshes + 1, 0);
      if (cfg.mode == 1 || cfg.mode == 2 || cfg.mode == 4) {
        const uint64_t budget = (uint64_t)cfg.lds_bytes;
        const bool split = cfg.mode != 1;
        const bool two = cfg.mode == 4;
        const bool nar = narrow_codes;
        const uint32_t gmax =
            split ? std::min(cfg.group_size, two ? kAxisSlots : 16) : cfg.group_size;
        // bytes of LDS: staged words + per-meshlet code storage ("codes"
        // counts vertices in mode 1, block words in mode 2, code words and
        // staged axis words in mode 4, whose axis pass stages nothing else)
        auto need = [split, two](uint64_t words, uint64_t cu) {
          return two ? 4 * cu
                     : (split ? 4 * ((words + 3) & ~3ull) + 4 * cu : 8 * ((words + 1) / 2) + 8 * cu);
        };
        group_lds = 0;
        for (int i = 0; i < n_meshes; ++i) {
          g_base[i] = (uint32_t)groups.size();
          const uint64_t bb = mp[i].blob_byte * 8ull;
          GroupEntry cur{};
          uint64_t codes = 0;
          bool open = false;
          auto close = [&]() {
            if (open) {
              // mode 1: packed codes after the words, in 8-byte units;
              // mode 2: walk/codes blocks after the words, in 4-byte units
              cur.codes_off = split ? ((cur.n_words + 3) & ~3u) : (cur.n_words + 1) / 2;
              groups.push_back(cur);
              group_lds = std::max<uint32_t>(group_lds, (uint32_t)need(cur.n_words, codes));
            }
            open = false;
          };
          for (uint32_t j = m_base[i]; j < m_base[i + 1]; ++j) {
            if (legacy[j]) { close(); continue; }
            const MeshletEntry& e = tab[j];
            const uint64_t lo = (bb + e.start_bit) >> 5;
            const uint64_t hi = ((bb + e.end_bit + 31) >> 5) + 2;
            const uint64_t unit = two ? split_code_words(e.n_local, nar) + axis_stage_words(bb, e)
                                      : (split ? split_block_words(e.n_local, nar) : e.n_local);
            if (open && (j - cur.m0 < gmax) && need(hi - cur.w_lo, codes + unit) <= budget) {
              cur.m1 = j + 1;
              cur.n_words = (uint32_t)(hi - cur.w_lo);
              codes += unit;
              continue;
            }
            close();
            if (need(hi - lo, unit) > budget) {   // too large for any group
              legacy[j] = 1;
              continue;
            }
            cur = GroupEntry{};
            cur.m0 = j;
            cur.m1 = j + 1;
            cur.w_lo = (uint32_t)lo;
            cur.n_words = (uint32_t)(hi - lo);
            codes = unit;
            open = true;
          }
          close();
        }
        g_base[n_meshes] = (uint32_t)groups.size();
      }
      n_groups = (uint32_t)groups.size();
      if (n_groups) {
        HIP_TRY(hipMalloc(&d_groups, sizeof(GroupEntry) * n_groups));
        HIP_TRY(hipMemcpy(d_groups, groups.data(), sizeof(GroupEntry) * n_groups,
                          hipMemcpyHostToDevice));
      }
      if (cfg.mode == 4) {
        int rc = token_layout(groups);
        if (rc) return rc;
      }
      if (pass == 1 || n_groups == 0) break;
      // validation decode of the groups
      uint32_t* d_flags = nullptr;
      HIP_TRY(hipMalloc(&d_flags, 4 * n_groups));
      hipError_t e = hipMemset(d_flags, 0, 4 * n_groups);
      GroupLaunch g = group_args(d_v_own, d_t_own, 0, n_groups, d_flags, 0, n_tgroups);
      g.queue = nullptr;   // validation: one workgroup per group
      g.grid = 0;
      if (e == hipSuccess) e = launch_group(g, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      std::vector<uint32_t> flags(n_groups, 0);
      if (e == hipSuccess) e = hipMemcpy(flags.data(), d_flags, 4 * n_groups, hipMemcpyDeviceToHost);
      (void)hipFree(d_flags);
      HIP_TRY(e);
      bool any = false;
      for (uint32_t gi = 0; gi < n_groups; ++gi)
        if (flags[gi]) {
          any = true;
          for (uint32_t j = groups[gi].m0; j < groups[gi].m1; ++j) legacy[j] = 1;
        }
      if (!any) break;
      (void)hipFree(d_groups);
      d_groups = nullptr;
    }
    std::vector<uint32_t> lst;
    l_base.assign(n_meshes + 1, 0);
    for (int i = 0; i < n_meshes; ++i) {
      l_base[i] = (uint32_t)lst.size();
      for (uint32_t j = m_base[i]; j < m_base[i + 1]; ++j)
        if (legacy[j]) lst.push_back(j);
    }
    l_base[n_meshes] = (uint32_t)lst.size();
    n_legacy = (uint32_t)lst.size();
    if (n_legacy) {
      HIP_TRY(hipMalloc(&d_legacy, 4 * n_legacy));
      HIP_TRY(hipMemcpy(d_legacy, lst.data(), 4 * n_legacy, hipMemcpyHostToDevice));
      // validation decode of the wave-per-meshlet list: codes must fit int32
      uint32_t* d_flag = nullptr;
      HIP_TRY(hipMalloc(&d_flag, 4));
      hipError_t e = hipMemset(d_flag, 0, 4);
      DecodeLaunch d = legacy_args(d_v_own, d_t_own, 0, n_legacy);
      d.overflow = d_flag;
      d.use_queue = 0;
      d.grid = 0;
      if (e == hipSuccess) e = launch_decode(d, cfg, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      uint32_t flag = 0;
      if (e == hipSuccess) e = hipMemcpy(&flag, d_flag, 4, hipMemcpyDeviceToHost);
      (void)hipFree(d_flag);
      HIP_TRY(e);
      if (flag)
        return fail(STRIDE_E_UNSUPPORTED, "decoded codes exceed 32 bits (unsupported by the GPU decoder)");
    }
    ready = true;
    return STRIDE_OK;
  }

  // mode 4 token groups: unions of consecutive split groups of one mesh, up
  // to tok_group meshlets and kTokenLds bytes of staged words
  static constexpr uint64_t kTokenLds = 32768;
  int token_layout(const std::vector<GroupEntry>& groups) {
    const int n_meshes = (int)mp.size();
    std::vector<GroupEntry> tg;
    tg_base.assign(n_meshes + 1, 0);
    uint32_t tl = 0;
    for (int i = 0; i < n_meshes; ++i) {
      tg_base[i] = (uint32_t)tg.size();
      GroupEntry cur{};
      bool open = false;
      for (uint32_t gi = g_base[i]; gi < g_base[i + 1]; ++gi) {
        const GroupEntry& G = groups[gi];
        const uint64_t words = (uint64_t)G.w_lo + G.n_words - cur.w_lo;
        if (open && G.m0 == cur.m1 && G.m1 - cur.m0 <= (uint32_t)cfg.tok_group &&
            4 * words <= kTokenLds) {
          cur.m1 = G.m1;
          cur.n_words = (uint32_t)words;
          continue;
        }
        if (open) tg.push_back(cur);
        cur = GroupEntry{};
        cur.m0 = G.m0;
        cur.m1 = G.m1;
        cur.w_lo = G.w_lo;
        cur.n_words = G.n_words;
        open = true;
      }
      if (open) tg.push_back(cur);
    }
    tg_base[n_meshes] = (uint32_t)tg.size();
    // staged words + the token lanes' seen sets, 8 words each at stride
    // kTokenSeenStride
    for (const GroupEntry& G : tg)
      tl = std::max(tl, 4 * (G.n_words + 7 * kTokenSeenStride + (G.m1 - G.m0)));
    (void)hipFree(d_chains);
    d_chains = nullptr;
    n_chains = 0;
    if (cfg.tok_chains > 0) {
      int rc = chain_layout(groups, tg, tl);
      if (rc) return rc;
    }
    (void)hipFree(d_tgroups);
    d_tgroups = nullptr;
    n_tgroups = (uint32_t)tg.size();
    tlds_bytes = tl;
    if (n_tgroups) {
      HIP_TRY(hipMalloc(&d_tgroups, sizeof(GroupEntry) * n_tgroups));
      HIP_TRY(hipMemcpy(d_tgroups, tg.data(), sizeof(GroupEntry) * n_tgroups, hipMemcpyHostToDevice));
    }
    // walks at 16-byte aligned slots (walk_slot): n_v + 4 words per meshlet,
    // + 8 for the 4-entry reads past the last one
    if (!d_walk) HIP_TRY(hipMalloc(&d_walk, 4 * (n_v + 4ull * n_m + 8)));
    return STRIDE_OK;
  }

  // Chain token waves: consecutive meshlets of one mesh (those the split
  // groups decode) cut into token chains (chain_tokens with k0 =
  // tok_chains), up to 64 chains and tok_group meshlets per wave and
  // kChainLds bytes of LDS (staged words, 16-bit token records, 16 bytes of
  // final FIFO per chain, 256 first-sighting words).  Replaces tg and tl;
  // the chain-start scan kernel fills d_chains.
  static constexpr uint64_t kChainLds = 49152;
  int chain_layout(const std::vector<GroupEntry>& groups, std::vector<GroupEntry>& tg,
                   uint32_t& tl) {
    const int n_meshes = (int)mp.size();
    const uint32_t k0 = (uint32_t)cfg.tok_chains;
    std::vector<uint32_t> cbase(n_m, 0xFFFFFFFFu);
    tg.clear();
    tl = 0;
    uint32_t nc = 0;
    for (int i = 0; i < n_meshes; ++i) {
      tg_base[i] = (uint32_t)tg.size();
      const uint64_t bb = mp[i].blob_byte * 8ull;
      GroupEntry cur{};
      uint64_t rec_words = 0;
      bool open = false;
      auto need = [](uint64_t words, uint64_t rw) { return 4 * (words + rw) + 16 * 64 + 4 * 256; };
      auto close = [&]() {
        if (open) {
          tg.push_back(cur);
          tl = std::max<uint32_t>(tl, (uint32_t)need(cur.n_words, rec_words));
        }
        open = false;
      };
      for (uint32_t gi = g_base[i]; gi < g_base[i + 1]; ++gi) {
        for (uint32_t j = groups[gi].m0; j < groups[gi].m1; ++j) {
          const MeshletEntry& e = tab[j];
          const uint32_t nt = meshlet_tokens(e);
          const uint32_t cnt = chain_count(nt, chain_tokens(nt, e.n_local, e.n_strips, k0));
          const uint64_t rw = ((nt + 1) & ~1u) / 2;
          // the token section only (the header and strips, up to the
          // residual section), + 2 words for the lanes' one-token
          // look-ahead window
          const uint64_t lo = (bb + e.start_bit) >> 5;
          const uint64_t hi = ((bb + e.resid_bit + 31) >> 5) + 2;
          if (open && j == cur.m1 && j - cur.m0 < (uint32_t)cfg.tok_group &&
              cur.pad[1] - cur.pad[0] + cnt <= 64 &&
              need(hi - cur.w_lo, rec_words + rw) <= kChainLds) {
            cur.m1 = j + 1;
            cur.n_words = (uint32_t)(hi - cur.w_lo);
            cur.pad[1] += cnt;
            rec_words += rw;
          } else {
            close();
            if (need(hi - lo, rw) > kChainLds)
              return fail(STRIDE_E_UNSUPPORTED, "meshlet too large for the chain token pass");
            cur = GroupEntry{};
            cur.m0 = j;
            cur.m1 = j + 1;
            cur.w_lo = (uint32_t)lo;
            cur.n_words = (uint32_t)(hi - lo);
            cur.pad[0] = nc;
            cur.pad[1] = nc + cnt;
            rec_words = rw;
            open = true;
          }
          cbase[j] = nc;
          nc += cnt;
        }
      }
      close();
    }
    tg_base[n_meshes] = (uint32_t)tg.size();
    n_chains = nc;
    if (nc == 0) return STRIDE_OK;
    uint32_t* d_cb = nullptr;
    HIP_TRY(hipMalloc(&d_chains, sizeof(ChainEntry) * nc));
    HIP_TRY(hipMalloc(&d_cb, 4ull * n_m));
    hipError_t e = hipMemcpy(d_cb, cbase.data(), 4ull * n_m, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_chain_scan(d_buf, d_tab, d_mesh, d_cb, n_m, k0, d_chains, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_cb);
    HIP_TRY(e);
    return STRIDE_OK;
  }

  // the scratch a decode's kernels use: the token pass's walk buffer and the
  // work-queue counter (the decoder's own buffers, or the library's static
  // scratch for stride_decode_noalloc)
  struct Scratch {
    uint32_t* walk;
    uint32_t* queue;
  };
  Scratch own_scratch() const { return Scratch{d_walk, d_queue}; }
  // words of scratch a decode takes: 64 for the queue counter, then the
  // walk slots of the two-pass decode (walk_slot: n_v + 4 per meshlet + 8)
  uint64_t scratch_words() const {
    return 64ull + (cfg.mode == 4 ? n_v + 4ull * n_m + 8ull : 0ull);
  }

  GroupLaunch group_args(float* V, uint32_t* T, uint32_t g0, uint32_t g1, uint32_t* flags,
                         uint32_t tg0 = 0, uint32_t tg1 = 0, const Scratch* sc = nullptr) const {
    const Scratch S = sc ? *sc : own_scratch();
    GroupLaunch g;
    g.buf = d_buf;
    g.tab = d_tab;
    g.meshes = d_mesh;
    g.groups = d_groups + g0;
    g.n_groups = g1 - g0;
    g.V = V;
    g.T = T;
    g.overflow = flags;
    g.lds_bytes = group_lds;   // the largest group's need (<= the lds_bytes budget)
    g.grid = (uint32_t)cfg.grid;
    g.split = cfg.mode == 2 || cfg.mode == 4;
    g.narrow = narrow_codes;
    g.queue = (cfg.mode == 2 && cfg.use_queue && cfg.grid > 0) ? S.queue : nullptr;
    const bool two = cfg.mode == 4;
    g.tgroups = two ? d_tgroups + tg0 : nullptr;
    g.n_tgroups = two ? tg1 - tg0 : 0;
    g.tlds_bytes = tlds_bytes;
    g.walk = two ? S.walk : nullptr;
    g.chains = two ? d_chains : nullptr;
    g.k0 = (uint32_t)cfg.tok_chains;
    if (two) g.grid = 0;
    return g;
  }

  // side tables a decode reads: meshlet and mesh tables, the groups of both
  // passes, the token chains and the wave-per-meshlet list
  uint64_t side_bytes() const {
    const bool two = cfg.mode == 4;
    return (uint64_t)n_m * sizeof(MeshletEntry) + (uint64_t)mp.size() * sizeof(MeshParams) +
           (uint64_t)n_groups * sizeof(GroupEntry) +
           (two ? (uint64_t)n_tgroups * sizeof(GroupEntry) + (uint64_t)n_chains * sizeof(ChainEntry)
                : 0ull) +
           4ull * n_legacy;
  }

  // every mesh has at most 16 bits per axis: the split kernels keep uint16 codes
  bool narrow() const {
    for (const MeshParams& p : mp)
      for (int d = 0; d < 3; ++d)
        if (p.g_bits[d] > 16) return false;
    return true;
  }

  DecodeLaunch legacy_args(float* V, uint32_t* T, uint32_t i0, uint32_t i1,
                           const Scratch* sc = nullptr) const {
    const Scratch S = sc ? *sc : own_scratch();
    DecodeLaunch d;
    d.buf = d_buf;
    d.tab = d_tab;
    d.meshes = d_mesh;
    d.idx = cfg.mode == 0 ? nullptr : d_legacy + i0;
    d.overflow = nullptr;
    d.n_m = i1 - i0;
    d.V = V;
    d.T = T;
    d.queue = S.queue;
    d.use_queue = cfg.use_queue;
    d.grid = (uint32_t)cfg.grid;
    if (cfg.mode == 0) {  // index range is a meshlet range
      d.tab = d_tab + i0;
    }
    return d;
  }

  // Enqueue the decode of meshes [mesh0, mesh1) (scratch: the decoder's own
  // buffers unless sc is given).
  int enqueue(float* V, uint32_t* T, hipStream_t s, int mesh0, int mesh1,
              const Scratch* sc = nullptr) {
    if (!ready)
      return fail(STRIDE_E_HIP, "the decoder has no launch layout (the last set_tune failed)");
    const Scratch S = sc ? *sc : own_scratch();
    if (cfg.mode == 1 || cfg.mode == 2 || cfg.mode == 4) {
      const uint32_t g0 = g_base[mesh0], g1 = g_base[mesh1];
      if (g1 > g0) {
        const bool two = cfg.mode == 4;
        GroupLaunch g = group_args(V, T, g0, g1, nullptr, two ? tg_base[mesh0] : 0,
                                   two ? tg_base[mesh1] : 0, &S);
        if (g.queue) {   // a fill kernel, not a memset, so hipGraph capture stays clean
          hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, S.queue, 0u);
          HIP_TRY(hipGetLastError());
        }
        HIP_TRY(launch_group(g, s));
      }
      const uint32_t l0 = l_base[mesh0], l1 = l_base[mesh1];
      if (l1 > l0) {
        DecodeConfig c = cfg;
        c.use_queue = 0;
        c.grid = 0;
        DecodeLaunch d = legacy_args(V, T, l0, l1, &S);
        d.use_queue = 0;
        d.grid = 0;
        HIP_TRY(launch_decode(d, c, s));
      }
      return STRIDE_OK;
    }
    const uint32_t m0 = m_base[mesh0], m1 = m_base[mesh1];
    if (m1 == m0) return STRIDE_OK;
    if (cfg.use_queue) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, S.queue, 0u);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(launch_decode(legacy_args(V, T, m0, m1, &S), cfg, s));
    return STRIDE_OK;
  }

  // an output may be null only when it has no elements (an empty stream)
  bool outputs_ok(const float* V, const uint32_t* T) const { return (V || !n_v) && (T || !n_t); }

  int decode(float* V, uint32_t* T, hipStream_t s) {
    if (!outputs_ok(V, T)) return fail(STRIDE_E_ARG, "null output pointer");
    return enqueue(V, T, s, 0, (int)mp.size());
  }

  int decode_per_mesh(float* V, uint32_t* T, hipStream_t s) {
    if (!outputs_ok(V, T)) return fail(STRIDE_E_ARG, "null output pointer");
    for (int i = 0; i < (int)mp.size(); ++i) {
      int rc = enqueue(V, T, s, i, i + 1);
      if (rc) return rc;
    }
    return STRIDE_OK;
  }

  int decode_graph(float* V, uint32_t* T, hipStream_t s) {
    if (!outputs_ok(V, T)) return fail(STRIDE_E_ARG, "null output pointer");
    auto key = std::make_tuple((void*)V, (void*)T, cfg.mode, cfg.waves, cfg.stage_words,
                               cfg.grid, cfg.use_queue, cfg.group_size,
                               (cfg.lds_bytes * 128 + cfg.tok_group) * 32 + cfg.tok_chains);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      if (!cap_stream) HIP_TRY(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(V, T, cap_stream, 0, (int)mp.size());
      hipError_t ec = hipStreamEndCapture(cap_stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIP_TRY(ec);
      hipGraphExec_t ge = nullptr;
      hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ei);
      it = graphs.emplace(key, ge).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, s));
    return STRIDE_OK;
  }

  int set_tune(const stride_tune* t) {
    if (!t) return fail(STRIDE_E_ARG, "null tune");
    bool ok_w = t->waves == 1 || t->waves == 2 || t->waves == 4;
    bool ok_s = t->stage_words == 256 || t->stage_words == 512 || t->stage_words == 1024;
    bool ok_m = t->mode == 0 || t->mode == 1 || t->mode == 2 || t->mode == 4;
    bool ok_g = t->group_size >= 1 &&
                t->group_size <= (t->mode == 2 ? 16 : (t->mode == 4 ? kAxisSlots : 64));
    bool ok_l = t->lds_bytes >= 8192 && t->lds_bytes <= 65536 && t->lds_bytes % 8 == 0;
    if (!ok_w || !ok_s || !ok_m || !ok_g || !ok_l || t->grid < 0 || (t->use_queue && t->grid == 0))
      return fail(STRIDE_E_ARG,
                  "tune: mode in {0,1,2,4}, waves in {1,2,4}, stage_words in {256,512,1024}, "
                  "group_size in [1,64] ([1,16] for mode 2, [1,21] for mode 4), lds_bytes in "
                  "[8192,65536] "
                  "(multiple of 8), grid >= 0, "
                  "use_queue needs a persistent grid");
    const int tok_chains = t->tok_chains < 0 ? cfg.tok_chains : t->tok_chains;
    if (tok_chains > kMaxChains)
      return fail(STRIDE_E_ARG, "tok_chains must be in [0, 16] (or -1 to keep)");
    const int tok_group = t->tok_group ? t->tok_group : (tok_chains > 0 ? 4 : 32);
    if (tok_group < 1 || tok_group > 64) return fail(STRIDE_E_ARG, "tok_group must be in [1, 64]");
    const bool relayout = t->mode != cfg.mode || t->group_size != cfg.group_size ||
                          t->lds_bytes != cfg.lds_bytes || tok_group != cfg.tok_group ||
                          tok_chains != cfg.tok_chains;
    cfg.tok_group = tok_group;
    cfg.tok_chains = tok_chains;
    cfg.waves = t->waves;
    cfg.stage_words = t->stage_words;
    cfg.grid = t->grid;
    cfg.use_queue = t->use_queue;
    cfg.mode = t->mode;
    cfg.group_size = t->group_size;
    cfg.lds_bytes = t->lds_bytes;
    drop_graphs();
    return relayout ? layout() : STRIDE_OK;
  }
  void get_tune(stride_tune* t) const {
    t->waves = cfg.waves;
    t->stage_words = cfg.stage_words;
    t->grid = cfg.grid;
    t->use_queue = cfg.use_queue;
    t->mode = cfg.mode;
    t->group_size = cfg.group_size;
    t->lds_bytes = cfg.lds_bytes;
    t->reserved = (int32_t)n_legacy;
    t->tok_group = cfg.tok_group;
    t->tok_chains = cfg.tok_chains;
  }
};

struct stride_decoder {
  DecodeSet set;
  std::unique_ptr<PdltSet> v5;   // set for PDLT v5 streams
};
struct stride_batch {
  DecodeSet set;
};

struct stride_encoder {
  Prep prep;
  std::vector<int64_t> codes;
  // PDLT v5 state (stride_encoder_v5_prepare)
  std::vector<double> vn5, rows5;
  V5Header h5;
  bool have_v5 = false;
  Residuals res;
  bool have_res = false;
  int32_t pred_n[9];
  int32_t pred_K[3];
  int n_threads = 1;
};

extern "C" {

int stride_last_error(char* buf, size_t cap) {
  if (buf && cap) {
    size_t n = std::min(cap - 1, g_err.size());
    std::memcpy(buf, g_err.data(), n);
    buf[n] = 0;
  }
  return (int)g_err.size();
}

int stride_init(int device) {
  HIP_TRY(hipSetDevice(device));
  uint32_t* sp = nullptr;   // the device's static scratch address, looked up once
  return scratch_ptr(device, &sp);
}

int stride_version(void) { return 1; }

int stride_debug_pyset_order(const int32_t* in, int32_t n, int32_t* out) {
  if (!in || !out || n < 0) return fail(STRIDE_E_ARG, "null argument");
  std::vector<int> v(in, in + n);
  for (int x : v)
    if (x < 0) return fail(STRIDE_E_ARG, "negative key");
  std::vector<int> o = pyset_order(v);
  for (size_t i = 0; i < o.size(); ++i) out[i] = o[i];
  return (int)o.size();
}

// ----------------------------------------------------------------- decoder
int stride_decoder_create(const uint8_t* data, size_t len, stride_decoder** out) {
  if (!out) return fail(STRIDE_E_ARG, "null out");
  *out = nullptr;
  if (!data && len) return fail(STRIDE_E_ARG, "null data");
  auto h = std::make_unique<stride_decoder>();
  if (len >= 4 && be32(data) == kPdltMagic) {
    h->v5 = std::make_unique<PdltSet>();
    int rc = h->v5->build(data, len);
    if (rc) return rc;
    *out = h.release();
    return STRIDE_OK;
  }
  const uint8_t* blobs[1] = {data};
  size_t lens[1] = {len};
  int rc = h->set.build(blobs, lens, 1);
  if (rc) return rc;
  *out = h.release();
  return STRIDE_OK;
}

void stride_decoder_destroy(stride_decoder* h) { delete h; }

int stride_decoder_outputs(stride_decoder* h, float** d_verts, uint32_t** d_tris) {
  if (!h || !d_verts || !d_tris) return fail(STRIDE_E_ARG, "null argument");
  *d_verts = h->v5 ? h->v5->d_v_own : h->set.d_v_own;
  *d_tris = h->v5 ? h->v5->d_t_own : h->set.d_t_own;
  return STRIDE_OK;
}

int stride_dprb_from_dpr4(const uint8_t* data, size_t len, uint8_t* out, size_t cap,
                          size_t* out_len) {
  if (!out_len || (!data && len)) return fail(STRIDE_E_ARG, "null argument");
  std::vector<uint8_t> blob;
  std::string err;
  int code = 0;
  if (dprb_from_dpr4(data, len, blob, err, code)) return fail(code, err);
  *out_len = blob.size();
  if (out) {
    if (cap < blob.size()) return fail(STRIDE_E_ARG, "output buffer too small");
    std::memcpy(out, blob.data(), blob.size());
  }
  return STRIDE_OK;
}

int stride_decoder_create_dprb(const uint8_t* blob, size_t len, stride_decoder** out) {
  if (!out) return fail(STRIDE_E_ARG, "null out");
  *out = nullptr;
  if (!blob && len) return fail(STRIDE_E_ARG, "null blob");
  std::vector<uint8_t> dpr4;
  DprbTables t;
  std::string err;
  int code = 0;
  if (dprb_to_dpr4(blob, len, dpr4, t, err, code)) return fail(code, err);
  auto h = std::make_unique<stride_decoder>();
  const uint8_t* blobs[1] = {dpr4.data()};
  size_t lens[1] = {dpr4.size()};
  int rc = h->set.build(blobs, lens, 1);
  if (rc) return rc;
  // the blob's tables must be exactly what the pre-pass derived
  const DecodeSet& s = h->set;
  if (s.n_m != t.n_m || s.n_v != t.n_v || s.n_t != t.n_t)
    return fail(STRIDE_E_CORRUPT, "DPRB counts disagree with the meshlets");
  for (uint32_t m = 0; m < s.n_m; ++m) {
    const MeshletEntry& e = s.tab[m];
    bool ok = e.start_bit - t.body_start_bit == t.rel_off[m] && e.n_local == t.n_local[m] &&
              e.n_tris == t.n_tris[m] && (e.n_strips & kStripMask) == t.n_strips[m] &&
              e.v_off == t.v_off[m] && e.t_off == t.t_off[m] && e.n_kind0 == t.n_kind0[m] &&
              e.resid_bit - t.body_start_bit == t.resid_off[m];
    for (int j = 1; j < 6 && ok; ++j)
      ok = (uint16_t)(e.sub_bit[j] - e.resid_bit) == t.subs[5 * m + j - 1];
    if (!ok) return fail(STRIDE_E_CORRUPT, "DPRB table disagrees with meshlet " + std::to_string(m));
  }
  *out = h.release();
  return STRIDE_OK;
}

int stride_sblb_from_pdlt(const uint8_t* data, size_t len, uint8_t* out, size_t cap,
                          size_t* out_len) {
  if (!out_len || (!data && len)) return fail(STRIDE_E_ARG, "null argument");
  std::vector<uint8_t> blob;
  std::string err;
  if (int rc = sblb_from_pdlt(data, len, blob, err)) return fail(rc, err);
  *out_len = blob.size();
  if (out) {
    if (cap < blob.size()) return fail(STRIDE_E_ARG, "output buffer too small");
    std::memcpy(out, blob.data(), blob.size());
  }
  return STRIDE_OK;
}

int stride_decoder_create_sblb(const uint8_t* blob, size_t len, stride_decoder** out) {
  if (!out) return fail(STRIDE_E_ARG, "null out");
  *out = nullptr;
  if (!blob && len) return fail(STRIDE_E_ARG, "null blob");
  std::string err;
  if (int rc = sblb_check(blob, len, err)) return fail(rc, err);
  const uint8_t* s = nullptr;
  size_t sl = 0;
  sblb_split(blob, len, s, sl, err);
  return stride_decoder_create(s, sl, out);
}

int stride_decoder_sizes(const stride_decoder* h, uint64_t* n_v, uint64_t* n_t,
                         uint32_t* n_meshlets) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) {
    if (n_v) *n_v = h->v5->H.n_v;
    if (n_t) *n_t = h->v5->H.n_t;
    if (n_meshlets) *n_meshlets = h->v5->H.n_m;
    return STRIDE_OK;
  }
  if (n_v) *n_v = h->set.n_v;
  if (n_t) *n_t = h->set.n_t;
  if (n_meshlets) *n_meshlets = h->set.n_m;
  return STRIDE_OK;
}

int stride_decode(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) return h->v5->decode(d_verts, d_tris, 1, (hipStream_t)stream);
  return h->set.decode(d_verts, d_tris, (hipStream_t)stream);
}

int stride_decode_noalloc(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) return fail(STRIDE_E_ARG, "stride_decode_noalloc decodes DPR4 streams only");
  if (!h->set.outputs_ok(d_verts, d_tris)) return fail(STRIDE_E_ARG, "null output pointer");
  DecodeSet& S = h->set;
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != S.device)
    return fail(STRIDE_E_ARG, "the decoder's device is not the current device");
  const uint64_t need = S.scratch_words();
  if (need > kScratchWords) {
    char b[200];
    std::snprintf(b, sizeof b,
                  "stride_decode_noalloc: this decode needs %llu words of scratch, the static "
                  "scratch holds %llu (32 MiB)",
                  (unsigned long long)need, (unsigned long long)kScratchWords);
    return fail(STRIDE_E_UNSUPPORTED, b);
  }
  uint32_t* sp = nullptr;
  int rc = scratch_ptr(dev, &sp);
  if (rc) return rc;
  const DecodeSet::Scratch sc{sp + 64, sp};
  return S.enqueue(d_verts, d_tris, (hipStream_t)stream, 0, (int)S.mp.size(), &sc);
}

int stride_decode_normalized(stride_decoder* h, float* d_verts, uint32_t* d_tris, void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (!h->v5) return fail(STRIDE_E_ARG, "normalised output is defined for PDLT v5 streams only");
  return h->v5->decode(d_verts, d_tris, 0, (hipStream_t)stream);
}

int stride_decoder_set_precision(stride_decoder* h, int32_t chain_bits) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (chain_bits != 64 && chain_bits != 32) return fail(STRIDE_E_ARG, "chain_bits must be 64 or 32");
  if (!h->v5) {
    if (chain_bits == 64) return STRIDE_OK;   // DPR4 decodes integer codes exactly
    return fail(STRIDE_E_ARG, "the float32 chain is defined for PDLT v5 streams only");
  }
  const int c32 = chain_bits == 32 ? 1 : 0;
  if (c32 == h->v5->chain32) return STRIDE_OK;
  // the chain's LDS block is three float32 or three float64 per vertex; a
  // failed layout leaves the previous one (and precision) in place
  const int prev = h->v5->chain32;
  h->v5->chain32 = c32;
  const int rc = h->v5->set_layout(h->v5->group_size, h->v5->lds_budget, h->v5->tok_lds_budget);
  if (rc) h->v5->chain32 = prev;
  return rc;
}

int stride_decoder_format(const stride_decoder* h, int32_t* version, float* lin3,
                          uint32_t* n_boundary) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (version) *version = h->v5 ? 5 : 3;
  if (lin3)
    for (int i = 0; i < 3; ++i) lin3[i] = h->v5 ? h->v5->H.lin3[i] : 0.f;
  if (n_boundary) *n_boundary = h->v5 ? h->v5->H.n_b : 0u;
  return STRIDE_OK;
}

int stride_decode_to_host(stride_decoder* h, float* verts, uint32_t* tris) {
  if (!h) return fail(STRIDE_E_ARG, "null argument");
  if (h->v5) {
    PdltSet& p = *h->v5;
    if ((!verts && p.H.n_v) || (!tris && p.H.n_t)) return fail(STRIDE_E_ARG, "null argument");
    int rc = p.own_outputs();
    if (rc) return rc;
    rc = p.decode(p.d_v_own, p.d_t_own, 1, nullptr);
    if (rc) return rc;
    if (p.H.n_v) HIP_TRY(hipMemcpy(verts, p.d_v_own, 12 * p.H.n_v, hipMemcpyDeviceToHost));
    if (p.H.n_t) HIP_TRY(hipMemcpy(tris, p.d_t_own, 12 * p.H.n_t, hipMemcpyDeviceToHost));
    return STRIDE_OK;
  }
  if ((!verts && h->set.n_v) || (!tris && h->set.n_t))
    return fail(STRIDE_E_ARG, "null argument");
  DecodeSet& s = h->set;
  int rc = s.decode(s.d_v_own, s.d_t_own, nullptr);
  if (rc) return rc;
  if (s.n_v) HIP_TRY(hipMemcpy(verts, s.d_v_own, 12 * s.n_v, hipMemcpyDeviceToHost));
  if (s.n_t) HIP_TRY(hipMemcpy(tris, s.d_t_own, 12 * s.n_t, hipMemcpyDeviceToHost));
  return STRIDE_OK;
}

int stride_decoder_bytes(const stride_decoder* h, uint64_t* stream_bytes, uint64_t* side_bytes) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) {
    if (stream_bytes) *stream_bytes = h->v5->H.stream_bytes;
    if (side_bytes) *side_bytes = h->v5->side_bytes();
    return STRIDE_OK;
  }
  if (stream_bytes) *stream_bytes = h->set.stream_bytes;
  if (side_bytes) *side_bytes = h->set.side_bytes();
  return STRIDE_OK;
}

int stride_decoder_header(const stride_decoder* h, float* f13, int32_t* i16) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) {
    const PdltHost& p = h->v5->H;
    if (f13) {
      for (int i = 0; i < 3; ++i) f13[i] = p.center[i];
      f13[3] = p.scale;
      f13[4] = p.pce;
      for (int i = 0; i < 3; ++i) f13[5 + i] = p.g_min[i];
      for (int i = 0; i < 3; ++i) f13[8 + i] = p.g_range[i];
      f13[11] = f13[12] = 0.f;
    }
    if (i16) {
      for (int i = 0; i < 16; ++i) i16[i] = 0;
      for (int i = 0; i < 3; ++i) i16[i] = p.g_bits[i];
      i16[15] = (int32_t)p.code_width;
    }
    return STRIDE_OK;
  }
  const DupHeader& d = h->set.hdr[0];
  if (f13) {
    for (int i = 0; i < 3; ++i) f13[i] = d.center[i];
    f13[3] = d.scale;
    f13[4] = d.pce;
    for (int i = 0; i < 3; ++i) f13[5 + i] = d.g_min[i];
    for (int i = 0; i < 3; ++i) f13[8 + i] = d.g_range[i];
    f13[11] = f13[12] = 0.f;
  }
  if (i16) {
    for (int i = 0; i < 3; ++i) i16[i] = d.g_bits[i];
    for (int i = 0; i < 9; ++i) i16[3 + i] = d.pred_n[i];
    for (int i = 0; i < 3; ++i) i16[12 + i] = d.pred_K[i];
    i16[15] = (int32_t)d.code_width;
  }
  return STRIDE_OK;
}

int stride_decoder_set_tune(stride_decoder* h, const stride_tune* t) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (h->v5) {
    if (!t) return fail(STRIDE_E_ARG, "null argument");
    if (t->mode != 0 && t->mode != 3 && t->mode != 4)
      return fail(STRIDE_E_ARG, "PDLT v5 modes: 3 (one kernel) or 4 (token pass + chain pass)");
    const int tok = t->stage_words ? 4 * t->stage_words : h->v5->tok_lds_budget;
    const int kc = t->tok_chains < 0 ? h->v5->tok_chains : t->tok_chains;
    if (kc < 0 || kc > kMaxChains) return fail(STRIDE_E_ARG, "tok_chains must be in [0, 16]");
    const int tg = t->tok_group ? t->tok_group : (kc > 0 ? 4 : 64);
    if (tg < 1 || tg > 64) return fail(STRIDE_E_ARG, "tok_group must be in [1, 64]");
    // the mode decides the chain groups' blocks (the one-kernel decode keeps
    // the refs in LDS), so it is set before the layout
    const int prev_tg = h->v5->tok_group, prev_kc = h->v5->tok_chains,
              prev_tp = h->v5->two_pass;
    h->v5->tok_group = tg;
    h->v5->tok_chains = kc;
    if (t->mode) h->v5->two_pass = t->mode == 4;
    const int rc = h->v5->set_layout(t->group_size, t->lds_bytes, tok);
    if (rc) {
      h->v5->tok_group = prev_tg;
      h->v5->tok_chains = prev_kc;
      h->v5->two_pass = prev_tp;
      return rc;
    }
    return STRIDE_OK;
  }
  return h->set.set_tune(t);
}
int stride_decoder_get_tune(const stride_decoder* h, stride_tune* t) {
  if (!h || !t) return fail(STRIDE_E_ARG, "null argument");
  if (h->v5) {
    std::memset(t, 0, sizeof *t);
    t->mode = h->v5->two_pass ? 4 : 3;
    t->waves = 1;
    t->group_size = h->v5->group_size;
    t->lds_bytes = h->v5->lds_budget;
    t->stage_words = h->v5->tok_lds_budget / 4;
    // token lanes in place of chains (a meshlet too large for a chain wave)
    // run 64 meshlets a wave
    const bool fell_back = h->v5->tok_chains > 0 && !h->v5->chain_waves;
    t->tok_group = fell_back ? 64 : h->v5->tok_group;
    t->tok_chains = h->v5->chain_waves ? h->v5->tok_chains : 0;
    return STRIDE_OK;
  }
  h->set.get_tune(t);
  return STRIDE_OK;
}

// ------------------------------------------------------------------- batch
int stride_batch_create(const uint8_t* const* blobs, const size_t* lens, int32_t n_meshes,
                        stride_batch** out) {
  if (!out || !blobs || !lens) return fail(STRIDE_E_ARG, "null argument");
  *out = nullptr;
  auto h = std::make_unique<stride_batch>();
  int rc = h->set.build(blobs, lens, n_meshes);
  if (rc) return rc;
  *out = h.release();
  return STRIDE_OK;
}
int stride_batch_create_tuned(const uint8_t* const* blobs, const size_t* lens, int32_t n_meshes,
                              const stride_tune* tune, stride_batch** out) {
  if (!out) return fail(STRIDE_E_ARG, "null out");
  *out = nullptr;
  stride_batch* h = nullptr;
  int rc = stride_batch_create(blobs, lens, n_meshes, &h);
  if (rc) return rc;
  if (tune && (rc = h->set.set_tune(tune))) {
    delete h;
    return rc;
  }
  *out = h;
  return STRIDE_OK;
}
void stride_batch_destroy(stride_batch* h) { delete h; }
int stride_batch_outputs(stride_batch* h, float** d_verts, uint32_t** d_tris) {
  if (!h || !d_verts || !d_tris) return fail(STRIDE_E_ARG, "null argument");
  *d_verts = h->set.d_v_own;
  *d_tris = h->set.d_t_own;
  return STRIDE_OK;
}
int stride_batch_sizes(const stride_batch* h, uint64_t* n_v, uint64_t* n_t, uint32_t* n_meshlets) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (n_v) *n_v = h->set.n_v;
  if (n_t) *n_t = h->set.n_t;
  if (n_meshlets) *n_meshlets = h->set.n_m;
  return STRIDE_OK;
}
int stride_batch_offsets(const stride_batch* h, uint64_t* v_base, uint64_t* t_base) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (v_base) std::copy(h->set.v_base.begin(), h->set.v_base.end(), v_base);
  if (t_base) std::copy(h->set.t_base.begin(), h->set.t_base.end(), t_base);
  return STRIDE_OK;
}
int stride_batch_bytes(const stride_batch* h, uint64_t* stream_bytes, uint64_t* side_bytes) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  if (stream_bytes) *stream_bytes = h->set.stream_bytes;
  if (side_bytes) *side_bytes = h->set.side_bytes();
  return STRIDE_OK;
}
int stride_batch_set_tune(stride_batch* h, const stride_tune* t) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  return h->set.set_tune(t);
}
int stride_batch_get_tune(const stride_batch* h, stride_tune* t) {
  if (!h || !t) return fail(STRIDE_E_ARG, "null argument");
  h->set.get_tune(t);
  return STRIDE_OK;
}
int stride_batch_decode(stride_batch* h, float* d_verts, uint32_t* d_tris, void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  return h->set.decode(d_verts, d_tris, (hipStream_t)stream);
}
int stride_batch_decode_graph(stride_batch* h, float* d_verts, uint32_t* d_tris, void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  return h->set.decode_graph(d_verts, d_tris, (hipStream_t)stream);
}
int stride_batch_decode_per_mesh(stride_batch* h, float* d_verts, uint32_t* d_tris,
                                 void* stream) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  return h->set.decode_per_mesh(d_verts, d_tris, (hipStream_t)stream);
}

// ----------------------------------------------------------------- encoder
int stride_encoder_create(const int64_t* tris, int64_t n_t, const double* vn, int64_t n_v,
                          const double* fn, const int64_t* codes, int32_t max_verts,
                          int32_t max_tris, const double* k6, const double* nm6,
                          const int32_t* seed_table, int32_t seed_rows, int32_t n_threads,
                          stride_encoder** out) {
  if (!out || !tris || !vn || !fn || !codes || !k6 || !nm6 || !seed_table)
    return fail(STRIDE_E_ARG, "null argument");
  *out = nullptr;
  if (n_t <= 0 || n_v <= 0) return fail(STRIDE_E_ARG, "empty mesh");
  PrepParams P;
  P.max_verts = max_verts;
  P.max_tris = max_tris;
  for (int i = 0; i < 6; ++i) {
    P.K[i] = k6[i];
    P.NM[i] = nm6[i];
  }
  P.seed_table = seed_table;
  P.seed_rows = seed_rows;
  P.n_threads = std::max(1, (int)n_threads);
  auto h = std::make_unique<stride_encoder>();
  h->n_threads = P.n_threads;
  try {
    if (!build_prep(tris, n_t, vn, n_v, fn, codes, P, h->prep))
      return fail(STRIDE_E_PREP, h->prep.error);
  } catch (const std::exception& e) {
    return fail(STRIDE_E_PREP, e.what());
  }
  h->codes.assign(codes, codes + 3 * n_v);
  *out = h.release();
  return STRIDE_OK;
}

void stride_encoder_destroy(stride_encoder* h) { delete h; }

int stride_encoder_counts(const stride_encoder* h, int64_t* n_meshlets, int64_t* n_local_total,
                          int64_t* n_tris_total, int64_t* n_strips_total,
                          int64_t* n_para_total) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  int64_t nl = 0, nt = 0, ns = 0, np = 0;
  for (size_t m = 0; m < h->prep.plans.size(); ++m) {
    const Plan& p = h->prep.plans[m];
    nl += (int64_t)p.l2g.size();
    nt += (int64_t)p.tris_local.size() / 3;
    ns += (int64_t)p.strip_len.size();
    for (const auto& w : h->prep.walks[m]) np += w.kind;
  }
  if (n_meshlets) *n_meshlets = (int64_t)h->prep.plans.size();
  if (n_local_total) *n_local_total = nl;
  if (n_tris_total) *n_tris_total = nt;
  if (n_strips_total) *n_strips_total = ns;
  if (n_para_total) *n_para_total = np;
  return STRIDE_OK;
}

int stride_encoder_export_plans(const stride_encoder* h, int32_t* n_local, int32_t* n_tris,
                                int32_t* n_strips, int32_t* n_bnd, int64_t* l2g,
                                uint8_t* tris_local, int32_t* strip_len, int32_t* strip_tris,
                                uint8_t* walk5) {
  if (!h) return fail(STRIDE_E_ARG, "null handle");
  size_t pl = 0, pt = 0, ps = 0, pst = 0, pw = 0;
  for (size_t m = 0; m < h->prep.plans.size(); ++m) {
    const Plan& p = h->prep.plans[m];
    const int nl = (int)p.l2g.size(), nt = (int)p.tris_local.size() / 3;
    if (n_local) n_local[m] = nl;
    if (n_tris) n_tris[m] = nt;
    if (n_strips) n_strips[m] = (int32_t)p.strip_len.size();
    if (n_bnd) n_bnd[m] = p.n_bnd;
    if (l2g) std::copy(p.l2g.begin(), p.l2g.end(), l2g + pl);
    if (tris_local) std::copy(p.tris_local.begin(), p.tris_local.end(), tris_local + 3 * pt);
    if (strip_len) std::copy(p.strip_len.begin(), p.strip_len.end(), strip_len + ps);
    if (strip_tris) std::copy(p.strip_tris.begin(), p.strip_tris.end(), strip_tris + pst);
    if (walk5)
      for (const auto& w : h->prep.walks[m]) {
        uint8_t* o = walk5 + 5 * pw++;
        o[0] = w.v; o[1] = w.kind; o[2] = w.a; o[3] = w.b; o[4] = w.c;
      }
    pl += nl;
    pt += nt;
    ps += p.strip_len.size();
    pst += p.strip_tris.size();
  }
  return STRIDE_OK;
}

int stride_encoder_para_samples(const stride_encoder* h, int32_t* A, int32_t* B, int32_t* C,
                                int32_t* V, int64_t* off) {
  if (!h || !A || !B || !C || !V || !off) return fail(STRIDE_E_ARG, "null argument");
  int64_t r = 0;
  off[0] = 0;
  const int64_t* cd = h->codes.data();
  for (size_t m = 0; m < h->prep.plans.size(); ++m) {
    const Plan& p = h->prep.plans[m];
    for (const auto& w : h->prep.walks[m]) {
      if (!w.kind) continue;
      for (int d = 0; d < 3; ++d) {
        A[3 * r + d] = (int32_t)cd[3 * p.l2g[w.a] + d];
        B[3 * r + d] = (int32_t)cd[3 * p.l2g[w.b] + d];
        C[3 * r + d] = (int32_t)cd[3 * p.l2g[w.c] + d];
        V[3 * r + d] = (int32_t)cd[3 * p.l2g[w.v] + d];
      }
      r++;
    }
    off[m + 1] = r;
  }
  return STRIDE_OK;
}

int stride_encoder_residuals(stride_encoder* h, const int32_t* pred_n, const int32_t* pred_K,
                             int64_t* n_values, int64_t* max_abs) {
  if (!h || !pred_n || !pred_K) return fail(STRIDE_E_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (pred_K[d] < 0 || pred_K[d] > 30) return fail(STRIDE_E_ARG, "predictor K outside [0, 30]");
  std::memcpy(h->pred_n, pred_n, sizeof h->pred_n);
  std::memcpy(h->pred_K, pred_K, sizeof h->pred_K);
  compute_residuals(h->prep, h->codes.data(), pred_n, pred_K, h->res, h->n_threads);
  h->have_res = true;
  if (n_values) *n_values = (int64_t)h->res.vals.size();
  if (max_abs) *max_abs = h->res.max_abs;
  return STRIDE_OK;
}

int stride_encoder_export_residuals(const stride_encoder* h, int32_t* vals, int64_t* stream_off) {
  if (!h || !h->have_res) return fail(STRIDE_E_ARG, "residuals not computed");
  const Residuals& R = h->res;
  const size_t n_m = h->prep.plans.size();
  if (R.max_abs > 0x7FFFFFFF) return fail(STRIDE_E_UNSUPPORTED, "residual exceeds int32");
  if (vals)
    for (size_t i = 0; i < R.vals.size(); ++i) vals[i] = (int32_t)R.vals[i];
  if (stream_off) {
    stream_off[0] = 0;
    for (size_t m = 0; m < n_m; ++m) {
      int64_t base = R.off[m];
      for (int s = 0; s < 6; ++s) {
        int64_t cnt = s < 3 ? R.n_delta[m] : R.n_para[m];
        stream_off[6 * m + s] = base;
        base += cnt;
      }
    }
    stream_off[6 * n_m] = R.off[n_m];
  }
  return STRIDE_OK;
}

static int check_selection(const stride_encoder* h, const int32_t* sel, int64_t n_sel) {
  if (!sel) return STRIDE_OK;
  const int64_t n_m = (int64_t)h->prep.plans.size();
  if (n_sel < 0) return fail(STRIDE_E_ARG, "negative selection size");
  for (int64_t i = 0; i < n_sel; ++i)
    if (sel[i] < 0 || sel[i] >= n_m) return fail(STRIDE_E_ARG, "meshlet index out of range");
  return STRIDE_OK;
}

int stride_encoder_features(const stride_encoder* h, int32_t source, const int32_t* sel,
                            int64_t n_sel, float* out) {
  if (!h || !out) return fail(STRIDE_E_ARG, "null argument");
  if (source != 0 && source != 1) return fail(STRIDE_E_ARG, "source must be 0 or 1");
  if (source == 1 && !h->have_res) return fail(STRIDE_E_ARG, "residuals not computed");
  if (int rc = check_selection(h, sel, n_sel)) return rc;
  meshlet_features(h->prep, h->codes.data(), source ? &h->res : nullptr, sel, n_sel, out,
                   h->n_threads);
  return STRIDE_OK;
}

int stride_encoder_para_columns(const stride_encoder* h, const int32_t* sel, int64_t n_sel,
                                int32_t* cols, int64_t* off) {
  if (!h || !off) return fail(STRIDE_E_ARG, "null argument");
  if (int rc = check_selection(h, sel, n_sel)) return rc;
  para_columns(h->prep, h->codes.data(), sel, n_sel, cols, off, h->n_threads);
  return STRIDE_OK;
}

int stride_encoder_write(const stride_encoder* h, const float* f13, const int32_t* g_bits,
                         const int32_t* pred_n, const int32_t* pred_K, const uint8_t* tagk,
                         uint8_t* out, size_t cap, size_t* len) {
  if (!h || !f13 || !g_bits || !pred_n || !pred_K || !tagk || !len)
    return fail(STRIDE_E_ARG, "null argument");
  if (!h->have_res) return fail(STRIDE_E_ARG, "residuals not computed");
  if (std::memcmp(pred_n, h->pred_n, sizeof h->pred_n) ||
      std::memcmp(pred_K, h->pred_K, sizeof h->pred_K))
    return fail(STRIDE_E_ARG, "predictor differs from the one the residuals were computed with");
  const size_t n_m = h->prep.plans.size();
  for (size_t m = 0; m < n_m; ++m)
    for (int s = 0; s < 6; ++s) {
      uint8_t tag = tagk[12 * m + 2 * s];
      if (tag != 1 && tag != 2) return fail(STRIDE_E_ARG, "coder tag must be 1 (Rice) or 2 (EG)");
    }
  Header H;
  for (int i = 0; i < 3; ++i) {
    H.center[i] = f13[i];
    H.g_min[i] = f13[5 + i];
    H.g_range[i] = f13[8 + i];
    H.g_bits[i] = g_bits[i];
    H.pred_K[i] = pred_K[i];
  }
  H.scale = f13[3];
  H.per_coord_err = f13[4];
  for (int i = 0; i < 9; ++i) H.pred_n[i] = pred_n[i];
  std::vector<uint8_t> bytes = write_dup(h->prep, h->res, H, tagk);
  *len = bytes.size();
  if (out) {
    if (cap < bytes.size()) return fail(STRIDE_E_ARG, "output buffer too small");
    std::memcpy(out, bytes.data(), bytes.size());
  }
  return STRIDE_OK;
}

// ------------------------------------------------------------ PDLT v5 writer
int stride_encoder_v5_prepare(stride_encoder* h, const double* vn, const float* f13,
                              const int32_t* g_bits, int64_t* n_boundary, int64_t* n_samples) {
  if (!h || !vn || !f13 || !g_bits) return fail(STRIDE_E_ARG, "null argument");
  V5Header H;
  for (int i = 0; i < 3; ++i) {
    H.center[i] = f13[i];
    H.g_min[i] = f13[5 + i];
    H.g_range[i] = f13[8 + i];
    H.g_bits[i] = g_bits[i];
    if (g_bits[i] < 1 || g_bits[i] > 30) return fail(STRIDE_E_ARG, "g_bits outside [1, 30]");
  }
  H.scale = f13[3];
  H.per_coord_err = f13[4];
  h->h5 = H;
  h->vn5.assign(vn, vn + 3 * h->prep.n_v);
  v5_samples(h->prep, h->codes.data(), h->vn5.data(), H, h->rows5, h->n_threads);
  h->have_v5 = true;
  if (n_boundary) *n_boundary = (int64_t)h->prep.boundary.size();
  if (n_samples) *n_samples = (int64_t)(h->rows5.size() / 12);
  return STRIDE_OK;
}

int stride_encoder_v5_samples(const stride_encoder* h, double* A, double* B, double* C,
                              double* T) {
  if (!h || !A || !B || !C || !T) return fail(STRIDE_E_ARG, "null argument");
  if (!h->have_v5) return fail(STRIDE_E_ARG, "stride_encoder_v5_prepare not called");
  const size_t n = h->rows5.size() / 12;
  double* dst[4] = {A, B, C, T};
  for (size_t i = 0; i < n; ++i)
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) dst[q][3 * i + d] = h->rows5[12 * i + 3 * q + d];
  return STRIDE_OK;
}

int stride_encoder_write_v5(const stride_encoder* h, const float* lin3, uint8_t* out, size_t cap,
                            size_t* len) {
  if (!h || !lin3 || !len) return fail(STRIDE_E_ARG, "null argument");
  if (!h->have_v5) return fail(STRIDE_E_ARG, "stride_encoder_v5_prepare not called");
  if (h->prep.boundary.size() > 0xFFFFFFFFull || h->prep.n_v > 0xFFFFFFFFll)
    return fail(STRIDE_E_ARG, "mesh too large for 32-bit PDLT fields");
  std::vector<uint8_t> bytes = write_v5(h->prep, h->codes.data(), h->vn5.data(), h->h5, lin3,
                                        h->n_threads);
  *len = bytes.size();
  if (out) {
    if (cap < bytes.size()) return fail(STRIDE_E_ARG, "output buffer too small");
    std::memcpy(out, bytes.data(), bytes.size());
  }
  return STRIDE_OK;
}

}  // extern "C"
