// Host-side launchers for the HIP kernels (one stream argument each; no
// allocation, no synchronisation).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gaussian_device.hpp"
#include "types.hpp"

namespace gs3d {

// ---- preprocess.hip --------------------------------------------------------

struct PreprocessArgs {
  const FrameConsts* fc;
  GaussianStreams g;
  uint32_t* indices;
  float* depth;
  uint32_t depth_len;
  Splat2D* splats;
  uint32_t* indirect_args;
  uint32_t* dispatch_args;
  uint32_t* status;
  uint32_t* ticket;
  uint32_t num_tiles;
  uint32_t* err;
  // preprocess.wesl:69-78: Gaussian i is skipped when bit i of `selection`
  // equals `invert_selection` (no selection: null)
  const uint32_t* selection;
  uint32_t invert_selection;
  // per Gaussian: the fragment depth of its splat (clip z / w of the centre,
  // render.wesl:108-125), for depth-tested draws (null: not written)
  float* splat_depth;
  // per Gaussian: the tile rectangle of its splat on rect_grid (tile_rect),
  // for the binning of the frame (null: not written)
  uint2* rects;
  TileGrid rect_grid;
  // the depth sort's first-pass histogram (digit 0 of the keys it will sort:
  // the visible keys and the 2.0 pads up to min(dispatch.x * 3840,
  // sort_max)), zeroed by k_frame_pre (null: the sort counts for itself)
  uint32_t* sort_hist;
  uint32_t sort_max;
  // thin splats whose axes k_thin_records redoes in the exact order: their
  // count (word kThinCountWord of the ticket buffer, zeroed by k_frame_pre)
  // and Gaussian indices (room for every Gaussian)
  uint32_t* thin_count;
  uint32_t* thin_idx;
};

// Word of the projection's ticket buffer (256 bytes) that counts the frame's
// thin splats: a cache line of its own, away from the ticket.
constexpr uint32_t kThinCountWord = 32;

uint32_t preprocess_num_tiles(uint32_t n);
void launch_frame_pre(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                      const GaussianTransformPod* gt, FrameConsts* fc, uint32_t* indirect_args,
                      uint32_t* dispatch_args, uint32_t* status, uint32_t num_tiles, uint32_t* ticket,
                      uint32_t* sort_hist = nullptr);
void launch_frame_consts(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                         const GaussianTransformPod* gt, FrameConsts* fc);
void launch_preprocess(hipStream_t st, int pod, const PreprocessArgs& a);
// splat_depth: as PreprocessArgs::splat_depth (or null)
void launch_vertex(hipStream_t st, int pod, const FrameConsts* fc, const GaussianStreams& g,
                   const uint32_t* indirect_args, const uint32_t* indices, Splat2D* splats, uint32_t max_v,
                   float* splat_depth);

// ---- radix_sort.hip --------------------------------------------------------

// Element count read on the device: n = min(*count * count_mul, count_max),
// or count_max when count is null (a count fixed at launch).
struct SortCount {
  const uint32_t* count;
  uint32_t count_mul;
  uint32_t count_max;
};

struct SortScratch {
  uint32_t* hist;      // [passes][256]
  uint32_t* lookback;  // [passes][max_tiles][256]
  uint32_t* tickets;   // [passes]
  uint32_t max_tiles;  // tiles of the smallest onesweep tile the scratch serves
  uint32_t* err;
  uint32_t items;  // keys per thread of a onesweep tile: 8, 12, 16, 20 or 24 (0: 20)
};

// Keys per thread a onesweep tile can have.
bool sort_items_supported(uint32_t items);
uint32_t sort_tile_items();
uint32_t sort_max_tiles(uint32_t count_max);
size_t sort_lookback_words(uint32_t count_max, int passes);
// Stable ascending LSD sort of (key, value) by key bits [0, 8*passes).
// Ping-pongs between (k0, v0) and (k1, v1); returns 0 when the result is in
// (k0, v0), 1 when it is in (k1, v1). Keys [pad_begin, pad_end) of
// pad_keys (when non-null) are set to 0xFFFFFFFF first.
// Every pass counts the next pass's digits as it goes; the first pass's
// histogram comes from an upsweep over the keys, or, with hist_ready, from
// whoever wrote the keys: then sc.hist[0..255] already holds the digit-0
// counts of the n keys and sc.hist[256..] is zero.
// k_first/v_first (when non-null): pass 0 reads them instead of (k0, v0),
// which it leaves unread; the passes after it alternate as above.
int launch_radix_sort(hipStream_t st, int num_cus, SortCount cnt, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                      uint32_t* v1, int passes, const SortScratch& sc, uint32_t* pad_keys = nullptr,
                      uint32_t pad_begin = 0, uint32_t pad_end = 0, bool hist_ready = false,
                      const uint32_t* k_first = nullptr, const uint32_t* v_first = nullptr);

// The allocation-free sort: a fixed device array per GPU as its scratch.
// noalloc_sort_init (with the device current, once per device) looks up the
// array's address on that device. launch_radix_sort_noalloc sorts n pairs from
// (keys_in, vals_in) into (keys_out, vals_out), stable and ascending, on
// stream st, allocating nothing; returns 1 when n exceeds
// GS3D_RADIX_SORT_NOALLOC_MAX_KEYS and 2 when the device was never
// initialised (nothing launched), 0 otherwise.
hipError_t noalloc_sort_init(int ordinal);
size_t noalloc_sort_scratch_bytes();
int launch_radix_sort_noalloc(hipStream_t st, int num_cus, int ordinal, const uint32_t* keys_in,
                              const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                              uint32_t* err);

// ---- tiles.hip ---------------------------------------------------------------

struct BinArgs {
  const uint32_t* indirect_args;  // instance count at [1]
  const uint32_t* indices;        // sorted Gaussian indices
  const Splat2D* splats;
  const uint2* rects;  // per-Gaussian tile rectangles on `grid` (null: from the splats)
  const FrameConsts* fc;  // max_std_dev, for rectangles made from the splats
  TileGrid grid;
  uint32_t* pair_keys;
  uint32_t* pair_vals;
  uint32_t capacity;
  uint32_t* pair_count;  // total pairs (may exceed capacity)
  uint32_t* status;
  uint32_t* ticket;
  uint32_t max_tiles;
  uint32_t* err;
  // the tile sort's first-pass histogram (digit 0 of the stored pairs' tile
  // ids), zeroed by k_bin_reset (null: the sort counts for itself)
  uint32_t* sort_hist;
  uint32_t num_gaussians;  // entries of splats and rects: indices at or above it are reported, not read
};

struct CompositeArgs {
  const FrameConsts* fc;
  const uint32_t* pair_vals;
  const uint32_t* ranges;  // [tiles][2]
  const Splat2D* splats;
  uint32_t num_gaussians;  // entries of splats and splat_depth: pair values at or above it are reported, not read
  uint32_t* err;
  TileGrid grid;
  int format;
  int blend;
  int load;  // blend over the target's contents instead of clearing to black
  uint32_t* frag_count;  // exact mode only: counts the non-discarded fragments (or null)
  void* out;
  // fast mode work list (see composite.hip)
  uint32_t capacity;   // pair capacity: bounds the work list
  uint32_t seg_len;    // splats per segment of a long tile list
  uint32_t* item_tile; // [composite_max_items]
  uint32_t* item_seg;  // [composite_max_items]
  uint32_t* seg_base;  // [tiles] first item of a multi-segment tile
  const uint32_t* pair_count;  // the frame's pair total: above capacity, nothing is drawn
  // depth test (Renderer::new's depth_stencil, src/renderer.rs:120-137):
  // per-Gaussian fragment depth, the W x H f32 depth attachment (null: no
  // test), the CompareFn and whether passing fragments write their depth
  const float* splat_depth;
  float* depth;
  int depth_cmp;
  int depth_write;
  uint32_t* plan;      // kPlanWords words: {items, multi-segment items}, the work ticket at kPlanTicketWord
  uint32_t* plan_counts;  // [composite_plan_blocks][2] per-block item counts
  float4* partial;     // [composite_max_heavy_items][tile pixels]
};

uint32_t bin_max_tiles(uint32_t max_instances);
// sort_hist: hist_words words zeroed as well (the tile sort's histograms)
void launch_bin_reset(hipStream_t st, uint32_t* ranges, uint32_t num_tiles, uint32_t* status, uint32_t max_bin_tiles,
                      uint32_t* ticket, uint32_t* pair_count, uint32_t* sort_hist = nullptr, uint32_t hist_words = 0);
void launch_bin(hipStream_t st, const BinArgs& a);
// pair_max: device word keeping the largest pair total since the host reset
// it; host_total: device alias of mapped host memory that receives that
// maximum (or null)
void launch_ranges(hipStream_t st, const uint32_t* pair_count, uint32_t capacity, const uint32_t* keys,
                   uint32_t* ranges, uint32_t num_tiles, int num_cus, uint32_t* pair_max, uint32_t* host_total,
                   uint32_t* err);
uint32_t composite_max_items(uint32_t num_tiles, uint32_t capacity, uint32_t seg_len);
uint32_t composite_max_heavy_items(uint32_t capacity, uint32_t seg_len);
uint32_t composite_plan_blocks(uint32_t num_tiles);
// words of CompositeArgs::plan; the work ticket sits in the second 128-byte
// line, apart from the counts every workgroup reads
constexpr uint32_t kPlanWords = 64;
constexpr uint32_t kPlanTicketWord = 32;
// main_begin/main_end, when not null, are recorded around the main
// compositing kernel (for per-kernel timing).
void launch_composite(hipStream_t st, const CompositeArgs& a, int num_cus, hipEvent_t main_begin = nullptr,
                      hipEvent_t main_end = nullptr);
void launch_clear(hipStream_t st, void* out, uint32_t pixels, int format);

}  // namespace gs3d
