// Many views and many models over device-resident Gaussians:
//   MultiModelViewer - the reference's MultiModelViewer (src/multi_model.rs:289-531):
//                      several models, one camera, drawn into one target in
//                      the caller's order;
//   BatchRenderer    - a batch of views (model, camera, resolution) rendered
//                      as one HIP graph spread over several queues, with a
//                      tile-size autotuner;
//   Server           - a continuous-batching request loop: requests arrive
//                      on a clock, are admitted as queues and HBM free up,
//                      overlap on the device, and report their latency.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "viewer.hpp"

namespace gs3d {

class MultiModelViewer {
 public:
  MultiModelViewer(Device& dev, TextureFormat format);
  // MultiModelViewer::insert_model (src/multi_model.rs:353-391); replaces a
  // model under the same key.
  void insert_model(const std::string& key, const gs3d_gaussian* g, uint32_t n, PodKind pod);
  // MultiModelViewer::remove_model (src/multi_model.rs:393-396); false when
  // the key is unknown.
  bool remove_model(const std::string& key);
  void update_camera_with_pod(const CameraPod& pod);
  void update_camera(const Camera& c, uint32_t w, uint32_t h) { update_camera_with_pod(camera_pod_new(c, w, h)); }
  // MultiModelViewer::update_model_transform (src/multi_model.rs:414-429):
  // ModelNotFound for an unknown key.
  void update_model_transform(const std::string& key, const float pos[3], const float rot[4], const float scale[3]);
  void update_gaussian_transform(float size, int32_t mode, uint32_t sh_deg, bool no_sh0, float max_std_dev);
  // MultiModelViewer::render (src/multi_model.rs:476-530): every listed model
  // is preprocessed and sorted, then drawn in the given order into one pass
  // (cleared to black first).
  void render(Texture& target, const std::vector<std::string>& keys);
  void set_blend(BlendMode b);
  size_t size() const { return models_.size(); }
  // The viewer of one model; ModelNotFound for an unknown key.
  Viewer& model(const std::string& key);

 private:
  Device& dev_;
  TextureFormat format_;
  std::map<std::string, std::unique_ptr<Viewer>> models_;
  CameraPod camera_{};
  bool have_camera_ = false;
  GaussianTransformPod gt_{1.0f, pack_gaussian_flags(kModeSplat, 3, false), 3.0f, 0u};
  BlendMode blend_ = kBlendExactUnorm8;
};

struct ViewSpec {
  uint32_t model;
  uint32_t width, height;
};

class BatchRenderer {
 public:
  // models: resident GaussiansBuffers shared by every view. queues: how many
  // device queues (streams) the batch's views are spread over.
  BatchRenderer(Device& dev, std::vector<std::shared_ptr<GaussiansBuffer>> models, int queues);
  ~BatchRenderer();
  // Allocates one view slot (viewer + target) per view. Drops any graph.
  void set_layout(const std::vector<ViewSpec>& views);
  // Picks, for every view, the fastest compositing tile (8x8, 16x16, 32x8)
  // by timing `reps` draws of each with the given cameras, then the fastest
  // sort tile (8, 12, 16, 20 or 24 keys per thread, for the depth and tile
  // sorts, the same for every view) by timing `reps` whole batches of each.
  void autotune(const CameraPod* pods, int reps);
  // Renders the batch. With use_graph, the first call records the whole
  // batch (every view's kernels, spread over the queues) as one HIP graph and
  // later calls replay it; the cameras are copied in before each launch.
  void render(const CameraPod* pods, bool use_graph);
  // Waits for the batch; re-renders views whose tile lists overflowed (and
  // drops the graph, whose buffers then changed).
  void finish();
  Texture& target(size_t i) { return *targets_.at(i); }
  Viewer& viewer(size_t i) { return *viewers_.at(i); }
  size_t views() const { return specs_.size(); }
  uint32_t tile_w(size_t i) const { return viewers_.at(i)->renderer.tile_w(); }

 private:
  void stage_cameras(const CameraPod* pods);
  void drop_graph();

  Device& dev_;
  std::vector<std::shared_ptr<GaussiansBuffer>> models_;
  std::vector<int> queues_;
  std::vector<ViewSpec> specs_;
  std::vector<std::unique_ptr<Viewer>> viewers_;
  std::vector<std::unique_ptr<Texture>> targets_;
  CameraPod* staging_ = nullptr;  // pinned, one pod per view
  size_t staging_len_ = 0;
  hipEvent_t staged_ = nullptr;   // the last copies out of staging_
  std::vector<hipEvent_t> fork_, join_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

struct ServeStats {
  uint64_t requests = 0;
  uint64_t completed = 0;
  double seconds = 0.0;
  double frames_per_s = 0.0;
  double p50_ms = 0.0, p99_ms = 0.0, mean_ms = 0.0, max_ms = 0.0;
  double mean_in_flight = 0.0;
  uint64_t peak_hbm_bytes = 0;
  uint64_t deferred_for_memory = 0;
};

struct ServeRequest {
  uint32_t model;
  uint32_t width, height;
  double arrival_s;  // on the server's clock, from the start of the run
  CameraPod pod;
};

class Server {
 public:
  // Requests are served on `queues` device queues; every admitted request
  // holds a viewer of its model and a target of its resolution, taken from
  // per-model / per-resolution pools that grow while the HBM budget allows.
  Server(Device& dev, std::vector<std::shared_ptr<GaussiansBuffer>> models, int queues, uint64_t hbm_budget);
  ~Server();
  // Serves the requests (sorted by arrival) in real time: a request is
  // admitted once it has arrived and a queue and enough of the budget are
  // free; its latency runs from its arrival to its frame's completion.
  // keep[i] != 0 copies request i's image into images[i] (RGBA8, w*h*4).
  ServeStats run(const std::vector<ServeRequest>& reqs, const std::vector<uint8_t*>* images = nullptr);
  // One request alone (nothing else in flight); image into out.
  void render_alone(const ServeRequest& r, uint8_t* out);

 private:
  struct Slot;
  Slot* acquire(const ServeRequest& r);
  void release(Slot* s);
  uint64_t footprint(uint32_t model, uint32_t w, uint32_t h) const;

  Device& dev_;
  std::vector<std::shared_ptr<GaussiansBuffer>> models_;
  std::vector<int> queues_;
  uint64_t budget_;
  uint64_t used_ = 0;
  std::vector<std::unique_ptr<Slot>> slots_;
};

}  // namespace gs3d
