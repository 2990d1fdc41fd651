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
};

uint32_t preprocess_num_tiles(uint32_t n);
void launch_frame_pre(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                      const GaussianTransformPod* gt, FrameConsts* fc, uint32_t* indirect_args,
                      uint32_t* dispatch_args, uint32_t* status, uint32_t num_tiles, uint32_t* ticket);
void launch_frame_consts(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                         const GaussianTransformPod* gt, FrameConsts* fc);
void launch_preprocess(hipStream_t st, int pod, const PreprocessArgs& a);
void launch_vertex(hipStream_t st, int pod, const FrameConsts* fc, const GaussianStreams& g,
                   const uint32_t* indirect_args, const uint32_t* indices, Splat2D* splats, uint32_t max_v);

// ---- radix_sort.hip --------------------------------------------------------

// Element count read on the device: n = min(*count * count_mul, count_max).
struct SortCount {
  const uint32_t* count;
  uint32_t count_mul;
  uint32_t count_max;
};

struct SortScratch {
  uint32_t* hist;      // [passes][256]
  uint32_t* lookback;  // [passes][max_tiles][256]
  uint32_t* tickets;   // [passes]
  uint32_t max_tiles;
  uint32_t* err;
};

uint32_t sort_tile_items();
uint32_t sort_max_tiles(uint32_t count_max);
size_t sort_lookback_words(uint32_t count_max, int passes);
// Stable ascending LSD sort of (key, value) by key bits [0, 8*passes).
// Ping-pongs between (k0, v0) and (k1, v1); returns 0 when the result is in
// (k0, v0), 1 when it is in (k1, v1). Keys [pad_begin, pad_end) of
// pad_keys (when non-null) are set to 0xFFFFFFFF first.
int launch_radix_sort(hipStream_t st, int num_cus, SortCount cnt, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                      uint32_t* v1, int passes, const SortScratch& sc, uint32_t* pad_keys = nullptr,
                      uint32_t pad_begin = 0, uint32_t pad_end = 0);

// ---- tiles.hip ---------------------------------------------------------------

struct TileGrid {
  uint32_t width, height;   // target pixels
  uint32_t tile_w, tile_h;  // tile pixels
  uint32_t tiles_x, tiles_y;
};

struct BinArgs {
  const uint32_t* indirect_args;  // instance count at [1]
  const uint32_t* indices;        // sorted Gaussian indices
  const Splat2D* splats;
  TileGrid grid;
  uint32_t* pair_keys;
  uint32_t* pair_vals;
  uint32_t capacity;
  uint32_t* pair_count;  // total pairs (may exceed capacity)
  uint32_t* status;
  uint32_t* ticket;
  uint32_t max_tiles;
  uint32_t* err;
};

struct CompositeArgs {
  const FrameConsts* fc;
  const uint32_t* pair_vals;
  const uint32_t* ranges;  // [tiles][2]
  const Splat2D* splats;
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
  uint32_t* plan;      // {items, multi-segment items, work ticket}
  uint32_t* plan_counts;  // [composite_plan_blocks][2] per-block item counts
  float4* partial;     // [composite_max_heavy_items][tile pixels]
};

uint32_t bin_max_tiles(uint32_t max_instances);
void launch_bin_reset(hipStream_t st, uint32_t* ranges, uint32_t num_tiles, uint32_t* status, uint32_t max_bin_tiles,
                      uint32_t* ticket, uint32_t* pair_count);
void launch_bin(hipStream_t st, const BinArgs& a);
// host_total: device alias of mapped host memory that receives the pair
// total (or null)
void launch_ranges(hipStream_t st, const uint32_t* pair_count, uint32_t capacity, const uint32_t* keys,
                   uint32_t* ranges, int num_cus, uint32_t* host_total);
uint32_t composite_max_items(uint32_t num_tiles, uint32_t capacity, uint32_t seg_len);
uint32_t composite_max_heavy_items(uint32_t capacity, uint32_t seg_len);
uint32_t composite_plan_blocks(uint32_t num_tiles);
// main_begin/main_end, when not null, are recorded around the main
// compositing kernel (for per-kernel timing).
void launch_composite(hipStream_t st, const CompositeArgs& a, int num_cus, hipEvent_t main_begin = nullptr,
                      hipEvent_t main_end = nullptr);
void launch_clear(hipStream_t st, void* out, uint32_t pixels, int format);

}  // namespace gs3d
