// Device-resident tables shared by the STRIDE decode kernels and their host
// launchers.  Plain structs, identical layout on host and device.
#pragma once

#include <cstdint>

namespace stride {

// Per-meshlet side table, derived from the DPR4 bytes once by the index
// pre-pass (the DPR4 stream stores only the meshlet start offsets).
// Bit offsets are relative to the start of the mesh's DPR4 blob.
struct MeshletEntry {
  uint32_t mesh;        // index into MeshParams
  uint32_t start_bit;   // meshlet header
  uint32_t v_off;       // first output vertex of this meshlet, within the mesh
  uint32_t t_off;       // first output triangle, within the mesh
  uint32_t n_local;
  uint32_t n_tris;
  uint32_t n_strips;
  uint32_t n_kind0;     // anchor + delta vertices
  uint32_t resid_bit;   // anchor start
  uint32_t sub_bit[6];  // starts of dx, dy, dz, px, py, pz
  uint32_t end_bit;     // one past the meshlet's last bit
};
static_assert(sizeof(MeshletEntry) == 64, "MeshletEntry is one 64-byte line");

// Per-mesh constants (DPR4 v3 global header) plus placement in the batch.
struct MeshParams {
  uint64_t blob_byte;   // byte offset of the blob in the device buffer (16-aligned)
  uint64_t v_base;      // first output vertex of the mesh
  uint64_t t_base;      // first output triangle of the mesh
  float g_min[3];
  float g_range[3];
  float center[3];
  float scale;
  int32_t g_bits[3];
  int32_t pred_n[9];    // sign-extended int16 numerators, row-major by axis
  int32_t pred_K[3];
  int32_t pad[1];
};
static_assert(sizeof(MeshParams) == 128, "MeshParams is 128 bytes");

// Index pre-pass output per meshlet (before prefix sums).
struct MeshletScan {
  uint32_t n_local, n_tris, n_strips, n_kind0;
  uint32_t resid_bit;
  uint32_t sub_bit[6];
  uint32_t end_bit;     // bit just after the last residual read
  uint32_t error;       // 0 = ok, see kScanErr*
  uint32_t flags;       // kFlag*
  uint32_t pad[2];
};
static_assert(sizeof(MeshletScan) == 64, "MeshletScan is 64 bytes");

enum : uint32_t {
  kScanOk = 0,
  kScanErrSize = 1,       // n_local/n_tris/n_strips outside supported range
  kScanErrOverrun = 2,    // a field reads past the meshlet's bit range
  kScanErrFifo = 3,       // FIFO index past the FIFO fill
  kScanErrIndex = 4,      // raw local id >= n_local
  kScanErrStrips = 5,     // strip lengths do not add up to n_tris
  kScanErrWalk = 6,       // walk does not cover every local vertex
  kScanErrTag = 7,        // residual coder parameter k > 24 (unsupported)
};

// MeshletEntry::n_strips high bits
constexpr uint32_t kFlagRawInFifo = 1u << 16;  // a raw token names an id already in the FIFO
constexpr uint32_t kStripMask = 0xFFFFu;

// A group of consecutive meshlets of one mesh decoded by one wave, one lane
// per meshlet; its bit range is staged in LDS as one contiguous word run.
struct GroupEntry {
  uint32_t m0, m1;        // meshlet range [m0, m1)
  uint32_t w_lo;          // first staged word, absolute in the device buffer
  uint32_t n_words;       // staged words
  uint32_t codes_off;     // LDS offset of the packed codes, in 8-byte units
  uint32_t pad[3];
};
static_assert(sizeof(GroupEntry) == 32, "GroupEntry is 32 bytes");

// A token chain: a run of consecutive strip tokens of one meshlet that one
// lane parses on its own (the chain token pass).  A meshlet's tokens are cut
// into up to kMaxChains chains; the chain-start scan records the parser state
// at each cut.  A chain after the first starts with placeholder FIFO entries
// (n_local + slot), resolved once the chain before it has finished.
// The entry holds only what the cut's position does not fix: a token wave
// finds a chain's meshlet from the chain-index-0 entries before it, and its
// last token from the meshlet's chain length (chain_end).
struct ChainEntry {
  uint32_t pos;     // bit offset from the meshlet header (low 16) | first token << 16
  uint32_t state;   // strip token index j (9 bits) | strip tokens lp2 << 9 (10 bits) |
                    // FIFO fill << 19 (5 bits) | chain index in the meshlet << 24 (4 bits)
};
static_assert(sizeof(ChainEntry) == 8, "ChainEntry is 8 bytes");
constexpr int kMaxChains = 16;     // chains per meshlet (4-bit chain index)
constexpr int kChainMaxLocal = 240;  // placeholders n_local + 0..15 must fit a byte

constexpr int kAxisSlots = 21;   // meshlets per wave of the two-pass axis pass (3 lanes each)

constexpr int kMaxLocal = 256;   // vertices per meshlet (max_verts cap)
constexpr int kMaxTris = 256;    // triangles per meshlet (max_tris cap)

}  // namespace stride
