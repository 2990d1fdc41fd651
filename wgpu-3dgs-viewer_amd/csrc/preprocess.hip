// Preprocessor kernels: the per-frame constants ("pre", preprocess.wesl:52-56)
// and the fused projection kernel ("main" + "post", preprocess.wesl:60-126,
// with the per-instance vertex-stage work of render.wesl:75-130 folded in).
//
// Projection is HBM-bound: every Gaussian's position record is streamed once
// (coalesced, 16 B/lane); the covariance stream is read only for off-screen
// centres and visible splats, the SH stream only for visible splats. Visible
// Gaussians are compacted in ascending index order with a wave ballot, a
// workgroup scan and a decoupled look-back across workgroups, so the output
// order is deterministic (the reference appends with an atomic counter,
// preprocess.wesl:101, whose order is not).
#include <hip/hip_runtime.h>

#include "device_util.hpp"
#include "gaussian_device.hpp"
#include "kernels.hpp"
#include "math.hpp"
#include "types.hpp"

namespace gs3d {

constexpr int kPreThreads = 256;
constexpr int kPreWaves = kPreThreads / kWave;
// Gaussians per thread; waves per SIMD the register budget is sized for
#ifndef GS3D_PRE_ITEMS
#define GS3D_PRE_ITEMS 4
#endif
#ifndef GS3D_PRE_WAVES_PER_EU
#define GS3D_PRE_WAVES_PER_EU 4
#endif
// 1: the next visible item's covariance and SH words are requested before
// the current record's geometry and stores; 0: after them
#ifndef GS3D_PRE_PIPELINE
#define GS3D_PRE_PIPELINE 1
#endif
constexpr int kPreItems = GS3D_PRE_ITEMS;
constexpr uint32_t kPreTile = kPreThreads * kPreItems;

__global__ __launch_bounds__(256) void k_frame_pre(const CameraPod* __restrict__ cam,
                                                   const ModelTransformPod* __restrict__ mt,
                                                   const GaussianTransformPod* __restrict__ gt,
                                                   FrameConsts* __restrict__ fc, uint32_t* indirect_args,
                                                   uint32_t* dispatch_args, uint32_t* status,
                                                   uint32_t num_tiles, uint32_t* ticket, uint32_t* sort_hist) {
  // the depth sort's histograms, which this frame's projection fills
  if (sort_hist)
    for (uint32_t j = threadIdx.x; j < 4u * 256u; j += blockDim.x) sort_hist[j] = 0u;
  if (threadIdx.x == 0) {
    FrameConsts c;
    derive_frame_consts(*cam, *mt, *gt, c);
    *fc = c;
    // preprocess.wesl:55 (instance_count = 0); an empty model leaves the
    // sort dispatch at zero workgroups (preprocess.wesl:114 with V = 0).
    indirect_args[1] = 0u;
    dispatch_args[0] = 0u;
    dispatch_args[1] = 1u;
    dispatch_args[2] = 1u;
    *ticket = 0u;
    ticket[kThinCountWord] = 0u;  // the thin-splat queue of this frame
  }
  for (uint32_t t = threadIdx.x; t < num_tiles; t += blockDim.x) status[t] = 0u;
}

// Frame constants only (a Renderer driven without the preprocessor).
__global__ __launch_bounds__(64) void k_frame_consts(const CameraPod* __restrict__ cam,
                                                     const ModelTransformPod* __restrict__ mt,
                                                     const GaussianTransformPod* __restrict__ gt,
                                                     FrameConsts* __restrict__ fc) {
  if (threadIdx.x == 0) {
    FrameConsts c;
    derive_frame_consts(*cam, *mt, *gt, c);
    *fc = c;
  }
}

// world = model * (pos, 1) and the centre's NDC ((proj * view) * world /
// w, preprocess.wesl:80-82) in the bit-exact order of math.hpp.
template <class F>
__device__ __forceinline__ void project_centre(const F& fc, float4 pc, float world[4], float& nx,
                                               float& ny, float& nz) {
  const float p4[4] = {pc.x, pc.y, pc.z, 1.0f};
  float clip[4];
  m4v(fc.model, p4, world);  // model_to_world
  m4v(fc.pv, world, clip);   // world_to_camera: (proj * view) * world
  nx = clip[0] / clip[3];
  ny = clip[1] / clip[3];
  nz = clip[2] / clip[3];
}

// preprocess.wesl:87-98: a culled centre is re-tested after moving it
// towards the screen by the projected major axis.
template <class F>
__device__ __forceinline__ bool bound_visible(const F& fc, const float pos[3], const float c[6], float nx,
                                              float ny, float nz) {
  const float sd = fc.max_std_dev;
  float ax[4];
  cov2d_axes(fc, pos, c, sd * fc.gsize, ax);
  const float mx = (ax[0] * sd) / fc.size[0];
  const float my = (ax[1] * sd) / fc.size[1];
  const float mlen = sqrtf(mx * mx + my * my);
  const float nlen = sqrtf((-nx) * (-nx) + (-ny) * (-ny));
  const float dx = (-nx) / nlen;
  const float dy = (-ny) / nlen;
  const float t = fminf(mlen, nlen);
  return !cull(nx + t * dx, ny + t * dy, nz);
}

// One workgroup per kPreTile Gaussians (taken by ticket, so that tiles finish
// roughly in order for the look-back). Three phases:
//   A  every item: centre NDC, depth key, centre cull;
//   B  the items whose centre is culled, gathered into a queue so that the
//      (expensive, rarely needed) bound re-test runs on full waves;
//   C  the visible count is published for the decoupled look-back, the
//      visible items' splat records are written, and only then does the
//      tile look back (its predecessors have mostly resolved meanwhile) and
//      write the visible indices and keys in Gaussian order.
template <int POD>
__global__ __launch_bounds__(kPreThreads) __attribute__((amdgpu_waves_per_eu(GS3D_PRE_WAVES_PER_EU))) void k_preprocess(
    PreprocessArgs a) {
  __shared__ uint32_t s_cnt[kPreItems][kPreWaves];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_base;
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_hist[256];  // digit 0 of this tile's depth keys
  __shared__ uint32_t s_qlen;
  __shared__ uint16_t s_q[kPreTile];    // slots (k * kPreThreads + tid) whose centre is culled
  __shared__ uint8_t s_qvis[kPreTile];  // their re-test, by slot
  if (a.sort_hist) s_hist[threadIdx.x] = 0u;
  if (threadIdx.x == 0) s_qlen = 0u;

  const ConstFrame& fc = const_frame(a.fc);
  const uint32_t tile = take_ticket(a.ticket, &s_tile);
  const uint32_t tid = threadIdx.x;
  const uint32_t w = tid / kWave;
  const uint32_t lane = tid % kWave;
  const uint32_t base = tile * kPreTile;

  bool vis[kPreItems];
  float key[kPreItems];
  uint64_t masks[kPreItems];

  // issue every item's position load first
  float4 pcs[kPreItems];
  bool act[kPreItems];
#pragma unroll
  for (int k = 0; k < kPreItems; ++k) {
    const uint32_t i = base + uint32_t(k) * kPreThreads + tid;
    act[k] = i < a.g.n;
    if (act[k] && a.selection) {
      // selection early-out (preprocess.wesl:69-78)
      const bool bit = ((a.selection[i >> 5] >> (i & 31u)) & 1u) != 0u;
      act[k] = bit != (a.invert_selection != 0u);
    }
    if (act[k]) pcs[k] = a.g.pc[i];
  }

  // A
  bool culled[kPreItems];
#pragma unroll
  for (int k = 0; k < kPreItems; ++k) {
    key[k] = 0.0f;
    culled[k] = false;
    vis[k] = false;
    if (act[k]) {
      float world[4], nx, ny, nz;
      project_centre(fc, pcs[k], world, nx, ny, nz);
      key[k] = 1.0f - nz;  // preprocess.wesl:105
      culled[k] = cull(nx, ny, nz);
      vis[k] = !culled[k];
    }
    const uint64_t m = __ballot(culled[k]);
    uint32_t qb = 0;
    if (lane == 0 && m) qb = atomicAdd(&s_qlen, uint32_t(__popcll(m)));
    qb = __shfl(qb, 0, kWave);
    if (culled[k]) s_q[qb + mask_rank(m)] = uint16_t(uint32_t(k) * kPreThreads + tid);
  }
  __syncthreads();

  // B
  const uint32_t qlen = s_qlen;
  for (uint32_t j = tid; j < qlen; j += kPreThreads) {
    const uint32_t e = s_q[j];
    const uint32_t i = base + e;
    const float4 pc = a.g.pc[i];
    float world[4], nx, ny, nz, c[6];
    project_centre(fc, pc, world, nx, ny, nz);
    cov3d_of<POD>(load_cov_words<POD>(a.g, i), c);
    const float pos[3] = {pc.x, pc.y, pc.z};
    s_qvis[e] = bound_visible(fc, pos, c, nx, ny, nz) ? 1u : 0u;
  }
  __syncthreads();

  // the tile's visible count is known: publish it now and look back only
  // after the splat records, so that predecessors have resolved meanwhile
#pragma unroll
  for (int k = 0; k < kPreItems; ++k) {
    if (culled[k]) vis[k] = s_qvis[uint32_t(k) * kPreThreads + tid] != 0u;
    masks[k] = __ballot(vis[k]);
    if (lane == 0) s_cnt[k][w] = uint32_t(__popcll(masks[k]));
  }
  __syncthreads();
  uint32_t run = 0;  // meaningful in thread 0
  if (tid == 0) {
    // exclusive scan over (item, wave) = ascending Gaussian index order
    for (int k = 0; k < kPreItems; ++k)
      for (int ww = 0; ww < kPreWaves; ++ww) {
        const uint32_t c = s_cnt[k][ww];
        s_cnt[k][ww] = run;
        run += c;
      }
    status_store(a.status + tile, (tile == 0 ? kFlagInclusive : kFlagAggregate) | run);
  }

  // C: the splat records of the visible items. An item's covariance and SH
  // words are requested as soon as the previous item's SH has been folded
  // into its colour, so that their latency overlaps the rest of that record
  // and its stores (requested after the stores, they would wait behind them).
  ShWords<POD> sw;
  CovWords cw;
  if (vis[0]) {
    cw = load_cov_words<POD>(a.g, base + tid);
    sh_load<POD>(a.g, base + tid, fc.sh_deg, sw);
  }
#pragma unroll
  for (int k = 0; k < kPreItems; ++k) {
    const uint32_t i = base + uint32_t(k) * kPreThreads + tid;
    const uint32_t cbits = __float_as_uint(pcs[k].w);
    float world[4], c[6], rgb[3];
    if (vis[k]) {
      const float p4[4] = {pcs[k].x, pcs[k].y, pcs[k].z, 1.0f};
      m4v(fc.model, p4, world);  // uncontracted: the fragment depth derives from it
      cov3d_of<POD, true>(cw, c);
      splat_colour<POD>(fc, sw, world, cbits, rgb);
    }
    if (GS3D_PRE_PIPELINE && k + 1 < kPreItems && vis[k + 1]) {
      cw = load_cov_words<POD>(a.g, i + kPreThreads);
      sh_load<POD>(a.g, i + kPreThreads, fc.sh_deg, sw);
    }
    bool thin = false;
    if (vis[k]) {
      // (pos and cw serve only an inline exact evaluation, which this kernel
      // leaves to k_thin_records: cw may already hold the next item's words)
      const float pos[3] = {pcs[k].x, pcs[k].y, pcs[k].z};
      thin = splat_geometry<POD, false>(fc, world, c, pos, cw, cbits, rgb, a.splats + i,
                                        a.splat_depth ? a.splat_depth + i : nullptr, a.rects ? a.rects + i : nullptr,
                                        &a.rect_grid);
    }
    if (thin) {
      // queued for k_thin_records (rare: one atomic per thin splat)
      const uint32_t q = atomicAdd(a.thin_count, 1u);
      a.thin_idx[q] = i;
    }
    if (!GS3D_PRE_PIPELINE && k + 1 < kPreItems && vis[k + 1]) {
      cw = load_cov_words<POD>(a.g, i + kPreThreads);
      sh_load<POD>(a.g, i + kPreThreads, fc.sh_deg, sw);
    }
  }

  if (w == 0) {
    // the first wave looks back (run is meaningful in thread 0)
    run = __shfl(run, 0, kWave);
    uint32_t excl = 0;
    if (tile > 0) {
      excl = lookback_exclusive_wave(a.status, tile, a.err);
      if (lane == 0) status_store(a.status + tile, kFlagInclusive | (excl + run));
    }
    if (lane == 0) {
      s_base = excl;
      s_total = excl + run;
      if (tile == a.num_tiles - 1) {
        // preprocess.wesl:108-116 (post): visible count and sort dispatch size
        const uint32_t v = excl + run;
        a.indirect_args[1] = v;
        a.dispatch_args[0] = (v + kHistoBlockKvs - 1u) / kHistoBlockKvs;
        a.dispatch_args[1] = 1u;
        a.dispatch_args[2] = 1u;
      }
    }
  }
  __syncthreads();

  const uint32_t out_base = s_base;
#pragma unroll
  for (int k = 0; k < kPreItems; ++k) {
    if (vis[k]) {
      const uint32_t i = base + uint32_t(k) * kPreThreads + tid;
      const uint32_t p = out_base + s_cnt[k][w] + mask_rank(masks[k]);
      a.indices[p] = i;   // preprocess.wesl:102
      a.depth[p] = key[k];  // preprocess.wesl:105
      if (a.sort_hist) atomicAdd(&s_hist[__float_as_uint(key[k]) & 0xFFu], 1u);
    }
  }
  if (a.sort_hist) {
    __syncthreads();
    if (s_hist[tid]) atomicAdd(&a.sort_hist[tid], s_hist[tid]);
    if (tile == a.num_tiles - 1 && tid == 0) {
      // the 2.0 pads in the sorted range (bits 0x40000000: digit 0 is 0)
      const uint32_t v = s_total;
      const uint32_t blocks = (v + kHistoBlockKvs - 1u) / kHistoBlockKvs;
      const uint64_t n64 = uint64_t(blocks) * kHistoBlockKvs;
      const uint32_t n = n64 < uint64_t(a.sort_max) ? uint32_t(n64) : a.sort_max;
      if (n > v) atomicAdd(&a.sort_hist[0], n - v);
    }
  }

  if (tile == a.num_tiles - 1) {
    // preprocess.wesl:118-125: pad keys up to the sort block size with 2.0
    const uint32_t v = s_total;
    const uint32_t padded = ((v + kHistoBlockKvs - 1u) / kHistoBlockKvs) * kHistoBlockKvs;
    const uint32_t end = padded < a.depth_len ? padded : a.depth_len;
    for (uint32_t j = v + tid; j < end; j += kPreThreads) a.depth[j] = __uint_as_float(kPadKeyBits);
  }
}

// Standalone vertex stage for a Renderer driven without the fused
// preprocessor: one thread per drawn instance (render.wesl:82-83).
template <int POD>
__global__ __launch_bounds__(256) void k_vertex(const FrameConsts* __restrict__ fcp, GaussianStreams g,
                                                const uint32_t* __restrict__ indirect_args,
                                                const uint32_t* __restrict__ indices, Splat2D* splats,
                                                float* splat_depth) {
  const ConstFrame& fc = const_frame(fcp);
  const uint32_t v = indirect_args[1];
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < v; s += gridDim.x * blockDim.x) {
    const uint32_t i = indices[s];
    if (i >= g.n) continue;
    const float4 pc = g.pc[i];
    const float p4[4] = {pc.x, pc.y, pc.z, 1.0f};
    float world[4], c[6];
    m4v(fc.model, p4, world);  // uncontracted: the fragment depth derives from it
    const CovWords cw = load_cov_words<POD>(g, i);
    cov3d_of<POD, true>(cw, c);
    const float pos[3] = {pc.x, pc.y, pc.z};  // thin splats' exact axes are computed inline here
    compute_splat<POD>(fc, g, i, __float_as_uint(pc.w), world, c, pos, cw, splats + i,
                       splat_depth ? splat_depth + i : nullptr);
  }
}

// The exact axes of the thin splats k_preprocess queued (their records were
// written with the contracted ones): a grid-stride loop over the queue.
template <int POD>
__global__ __launch_bounds__(256) void k_thin_records(const FrameConsts* __restrict__ fcp, GaussianStreams g,
                                                      const uint32_t* __restrict__ thin_count,
                                                      const uint32_t* __restrict__ thin_idx, Splat2D* splats,
                                                      uint2* rects, TileGrid rect_grid) {
  const ConstFrame& fc = const_frame(fcp);
  const uint32_t n = *thin_count;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t i = thin_idx[j];
    if (i >= g.n) continue;
    const float4 pc = g.pc[i];
    const float pos[3] = {pc.x, pc.y, pc.z};
    thin_splat_record<POD>(fc, pos, load_cov_words<POD>(g, i), splats + i, rects ? rects + i : nullptr, &rect_grid);
  }
}

uint32_t preprocess_num_tiles(uint32_t n) { return (n + kPreTile - 1) / kPreTile; }

void launch_frame_pre(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                      const GaussianTransformPod* gt, FrameConsts* fc, uint32_t* indirect_args,
                      uint32_t* dispatch_args, uint32_t* status, uint32_t num_tiles, uint32_t* ticket,
                      uint32_t* sort_hist) {
  hipLaunchKernelGGL(k_frame_pre, dim3(1), dim3(256), 0, st, cam, mt, gt, fc, indirect_args, dispatch_args,
                     status, num_tiles, ticket, sort_hist);
}

void launch_frame_consts(hipStream_t st, const CameraPod* cam, const ModelTransformPod* mt,
                         const GaussianTransformPod* gt, FrameConsts* fc) {
  hipLaunchKernelGGL(k_frame_consts, dim3(1), dim3(64), 0, st, cam, mt, gt, fc);
}

void launch_preprocess(hipStream_t st, int pod, const PreprocessArgs& a) {
  if (a.num_tiles == 0) return;
  if (pod == kPodSingle)
    hipLaunchKernelGGL(k_preprocess<kPodSingle>, dim3(a.num_tiles), dim3(kPreThreads), 0, st, a);
  else
    hipLaunchKernelGGL(k_preprocess<kPodHalf>, dim3(a.num_tiles), dim3(kPreThreads), 0, st, a);
  // thin splats are few: 64 workgroups loop over the queue
  if (pod == kPodSingle)
    hipLaunchKernelGGL(k_thin_records<kPodSingle>, dim3(64), dim3(256), 0, st, a.fc, a.g, a.thin_count, a.thin_idx,
                       a.splats, a.rects, a.rect_grid);
  else
    hipLaunchKernelGGL(k_thin_records<kPodHalf>, dim3(64), dim3(256), 0, st, a.fc, a.g, a.thin_count, a.thin_idx,
                       a.splats, a.rects, a.rect_grid);
}

void launch_vertex(hipStream_t st, int pod, const FrameConsts* fc, const GaussianStreams& g,
                   const uint32_t* indirect_args, const uint32_t* indices, Splat2D* splats, uint32_t max_v,
                   float* splat_depth) {
  if (max_v == 0) return;
  uint32_t blocks = (max_v + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (pod == kPodSingle)
    hipLaunchKernelGGL(k_vertex<kPodSingle>, dim3(blocks), dim3(256), 0, st, fc, g, indirect_args, indices, splats,
                       splat_depth);
  else
    hipLaunchKernelGGL(k_vertex<kPodHalf>, dim3(blocks), dim3(256), 0, st, fc, g, indirect_args, indices, splats,
                       splat_depth);
}

}  // namespace gs3d
