// MultiModelViewer, BatchRenderer and Server (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <thread>

namespace gs3d {

// ---- MultiModelViewer ----------------------------------------------------------

MultiModelViewer::MultiModelViewer(Device& dev, TextureFormat format) : dev_(dev), format_(format) {
  if (format != kFormatRgba8Unorm && format != kFormatRgba32Float)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "unknown texture format");
}

static void apply_gaussian_transform(Viewer& v, const GaussianTransformPod& gt) {
  v.update_gaussian_transform(gt.size, int32_t(gt.flags & 3u), (gt.flags >> 2) & 3u, ((gt.flags >> 4) & 1u) != 0,
                              gt.max_std_dev);
}

void MultiModelViewer::insert_model(const std::string& key, const gs3d_gaussian* g, uint32_t n, PodKind pod) {
  auto v = std::make_unique<Viewer>(dev_, format_, g, n, pod);
  if (have_camera_) v->update_camera_with_pod(camera_);
  apply_gaussian_transform(*v, gt_);
  v->renderer.set_blend(blend_);
  models_[key] = std::move(v);
}

bool MultiModelViewer::remove_model(const std::string& key) {
  auto it = models_.find(key);
  if (it == models_.end()) return false;
  dev_.wait();
  models_.erase(it);
  return true;
}

void MultiModelViewer::update_camera_with_pod(const CameraPod& pod) {
  camera_ = pod;
  have_camera_ = true;
  for (auto& kv : models_) kv.second->update_camera_with_pod(pod);
}

Viewer& MultiModelViewer::model(const std::string& key) {
  auto it = models_.find(key);
  if (it == models_.end()) throw Error(GS3D_ERROR_MODEL_NOT_FOUND, "model with the given key does not exist");
  return *it->second;
}

void MultiModelViewer::update_model_transform(const std::string& key, const float pos[3], const float rot[4],
                                              const float scale[3]) {
  auto it = models_.find(key);
  if (it == models_.end()) throw Error(GS3D_ERROR_MODEL_NOT_FOUND, "model with the given key does not exist");
  it->second->update_model_transform(pos, rot, scale);
}

void MultiModelViewer::update_gaussian_transform(float size, int32_t mode, uint32_t sh_deg, bool no_sh0,
                                                 float max_std_dev) {
  GaussianTransformPod gt{};
  gt.size = size;
  gt.flags = pack_gaussian_flags(mode, sh_deg, no_sh0);
  gt.max_std_dev = max_std_dev;
  for (auto& kv : models_) apply_gaussian_transform(*kv.second, gt);  // validates
  gt_ = gt;
}

void MultiModelViewer::set_blend(BlendMode b) {
  blend_ = b;
  for (auto& kv : models_) kv.second->renderer.set_blend(b);
}

void MultiModelViewer::render(Texture& target, const std::vector<std::string>& keys) {
  if (target.format() != format_)
    throw Error(GS3D_ERROR_INVALID_ARGUMENT, "render target format differs from the viewer's format");
  std::vector<Viewer*> vs;
  for (const std::string& k : keys) {
    auto it = models_.find(k);
    if (it == models_.end()) throw Error(GS3D_ERROR_MODEL_NOT_FOUND, "model with the given key does not exist");
    vs.push_back(it->second.get());
  }
  if (vs.empty()) {
    launch_clear(dev_.stream(0), target.dev(), target.width() * target.height(), target.format());
    return;
  }
  // src/multi_model.rs:491-503: every model is preprocessed and sorted ...
  for (Viewer* v : vs) {
    v->splats_buffer.set_rect_grid(v->renderer.grid_for(target));
    v->preprocessor.set_sort_hist(v->radix_sorter.hist(), v->radix_sorter.len());
    v->preprocessor.preprocess(v->gaussians_buffer.len());
    v->radix_sorter.sort(v->radix_sort_indirect_args_buffer, true);
  }
  // ... then drawn in the caller's order into one pass (:505-527)
  for (size_t i = 0; i < vs.size(); ++i) vs[i]->renderer.render(target, vs[i]->indirect_args_buffer, false, i == 0);

  // kept for Device::poll, with every listed model's state
  RecordedUnit u;
  std::vector<Viewer::State> states;
  for (Viewer* v : vs) {
    states.push_back(v->state());
    u.renderers.push_back(&v->renderer);
  }
  Texture* t = &target;
  u.replay = [this, t, keys, vs, states] {
    for (size_t i = 0; i < vs.size(); ++i) vs[i]->restore(states[i]);
    render(*t, keys);
  };
  u.textures = {&target};
  dev_.record_unit(std::move(u));
}

// ---- BatchRenderer ----------------------------------------------------------------

BatchRenderer::BatchRenderer(Device& dev, std::vector<std::shared_ptr<GaussiansBuffer>> models, int queues)
    : dev_(dev), models_(std::move(models)) {
  if (queues < 1) queues = 1;
  queues_.push_back(0);
  for (int q = 1; q < queues; ++q) queues_.push_back(dev.add_queue());
  fork_.resize(queues_.size());
  join_.resize(queues_.size());
  for (size_t q = 0; q < queues_.size(); ++q) {
    hip_check(hipEventCreateWithFlags(&fork_[q], hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&join_[q], hipEventDisableTiming), "hipEventCreate");
  }
  hip_check(hipEventCreateWithFlags(&staged_, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventRecord(staged_, dev.stream(0)), "hipEventRecord");
}

BatchRenderer::~BatchRenderer() {
  try {
    dev_.wait();
  } catch (...) {
  }
  drop_graph();
  for (size_t q = 0; q < fork_.size(); ++q) {
    hipEventDestroy(fork_[q]);
    hipEventDestroy(join_[q]);
  }
  if (staged_) hipEventDestroy(staged_);
  if (staging_) hipHostFree(staging_);
}

void BatchRenderer::drop_graph() {
  if (exec_) hipGraphExecDestroy(exec_);
  if (graph_) hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void BatchRenderer::set_layout(const std::vector<ViewSpec>& views) {
  dev_.wait();
  drop_graph();
  viewers_.clear();
  targets_.clear();
  specs_ = views;
  for (size_t i = 0; i < views.size(); ++i) {
    const ViewSpec& s = views[i];
    if (s.model >= models_.size()) throw Error(GS3D_ERROR_MODEL_NOT_FOUND, "view names an unknown model");
    const int q = queues_[i % queues_.size()];
    viewers_.push_back(std::make_unique<Viewer>(dev_, kFormatRgba8Unorm, models_[s.model], q));
    targets_.push_back(std::make_unique<Texture>(dev_, s.width, s.height, kFormatRgba8Unorm));
  }
  if (staging_len_ < views.size()) {
    if (staging_) hipHostFree(staging_);
    staging_ = nullptr;
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&staging_), std::max<size_t>(views.size(), 1) * sizeof(CameraPod),
                            hipHostMallocDefault),
              "hipHostMalloc");
    staging_len_ = views.size();
  }
}

void BatchRenderer::stage_cameras(const CameraPod* pods) {
  // the previous batch's copies out of the pinned staging have finished
  hip_check(hipEventSynchronize(staged_), "hipEventSynchronize");
  std::memcpy(staging_, pods, specs_.size() * sizeof(CameraPod));
}

void BatchRenderer::render(const CameraPod* pods, bool use_graph) {
  if (specs_.empty()) return;
  hipStream_t s0 = dev_.stream(0);
  // cameras first, in queue 0's order (the graph is launched on queue 0)
  stage_cameras(pods);
  for (size_t i = 0; i < specs_.size(); ++i)
    hip_check(hipMemcpyAsync(const_cast<CameraPod*>(viewers_[i]->camera_buffer.dev()), staging_ + i, sizeof(CameraPod),
                             hipMemcpyHostToDevice, s0),
              "hipMemcpyAsync(camera)");
  hip_check(hipEventRecord(staged_, s0), "hipEventRecord");

  auto enqueue_views = [&]() {
    // fork: every queue starts after queue 0's copies; join: queue 0 waits
    // for every queue
    hip_check(hipEventRecord(fork_[0], s0), "hipEventRecord");
    for (size_t q = 1; q < queues_.size(); ++q)
      hip_check(hipStreamWaitEvent(dev_.stream(queues_[q]), fork_[0], 0), "hipStreamWaitEvent");
    for (size_t i = 0; i < specs_.size(); ++i) viewers_[i]->render(*targets_[i], false);
    for (size_t q = 1; q < queues_.size(); ++q) {
      hip_check(hipEventRecord(join_[q], dev_.stream(queues_[q])), "hipEventRecord");
      hip_check(hipStreamWaitEvent(s0, join_[q], 0), "hipStreamWaitEvent");
    }
  };

  if (!use_graph) {
    enqueue_views();
    return;
  }
  if (!exec_) {
    // one eager pass sizes every buffer (nothing may allocate under capture)
    enqueue_views();
    finish();
    hip_check(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    enqueue_views();
    hip_check(hipStreamEndCapture(s0, &graph_), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  hip_check(hipGraphLaunch(exec_, s0), "hipGraphLaunch");
}

void BatchRenderer::finish() {
  dev_.wait();
  const uint32_t e = dev_.take_errors();
  throw_kernel_errors(e);
  // every view's viewer drew once since the last finish: a view whose lists
  // overflowed is drawn again with larger lists until it fits
  bool regrown = false;
  for (int attempt = 0;; ++attempt) {
    bool any = false;
    for (auto& v : viewers_)
      if (v->renderer.grow_if_overflowed()) {
        v->renderer.rerender();
        any = true;
      }
    if (!any) break;
    regrown = true;
    if (attempt >= 3) throw Error(GS3D_ERROR_KERNEL, "tile lists still overflowed after growing them three times");
    dev_.wait();
  }
  if (regrown) drop_graph();  // the lists moved: record again on the next graph render
}

// Milliseconds of `reps` calls of `work` on stream st, timed with HIP events.
template <class F>
static float time_on_stream(hipStream_t st, int reps, F&& work) {
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  hip_check(hipEventRecord(a, st), "hipEventRecord");
  for (int r = 0; r < reps; ++r) work();
  hip_check(hipEventRecord(b, st), "hipEventRecord");
  hip_check(hipEventSynchronize(b), "hipEventSynchronize");
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms;
}

void BatchRenderer::autotune(const CameraPod* pods, int reps) {
  if (reps < 1) reps = 1;
  static const uint32_t sizes[3][2] = {{8, 8}, {16, 16}, {32, 8}};
  static const uint32_t items[5] = {8, 12, 16, 20, 24};
  drop_graph();
  for (size_t i = 0; i < viewers_.size(); ++i) {
    Viewer& v = *viewers_[i];
    Texture& t = *targets_[i];
    v.update_camera_with_pod(pods[i]);
    hipStream_t st = dev_.stream(v.queue);
    // the compositing tile, timed over the draws (binning, tile sort,
    // compositing) of one projected and sorted frame
    float best = 1e30f;
    int best_k = 1;
    for (int k = 0; k < 3; ++k) {
      v.renderer.set_tile_size(sizes[k][0], sizes[k][1]);
      // warm: sizes the buffers (and the tile lists) for this tile size, so
      // that the timed draws below are complete frames
      v.render(t, false);
      dev_.wait();
      v.renderer.settle();
      const float ms = time_on_stream(st, reps, [&] { v.renderer.render(t, v.indirect_args_buffer, false); });
      if (ms < best) {
        best = ms;
        best_k = k;
      }
    }
    v.renderer.set_tile_size(sizes[best_k][0], sizes[best_k][1]);
  }
  // the sort tile (both sorts of every view), timed over whole batches: the
  // views run concurrently, where a tile that needs more registers (fewer
  // resident workgroups) loses more than it does in a view timed alone
  hipStream_t s0 = dev_.stream(0);
  float best = 1e30f;
  int best_k = 3;
  for (int k = 0; k < 5; ++k) {
    for (auto& v : viewers_) v->set_sort_items(items[k]);
    render(pods, false);
    finish();
    const float ms = time_on_stream(s0, reps, [&] { render(pods, false); });
    finish();
    if (ms < best) {
      best = ms;
      best_k = k;
    }
  }
  for (auto& v : viewers_) v->set_sort_items(items[best_k]);
  finish();
}

// ---- Server -----------------------------------------------------------------------

struct Server::Slot {
  uint32_t model = 0, width = 0, height = 0;
  std::unique_ptr<Viewer> viewer;
  std::unique_ptr<Texture> target;
  hipEvent_t done = nullptr;
  bool busy = false;
  uint64_t bytes = 0;
  size_t request = 0;
};

Server::Server(Device& dev, std::vector<std::shared_ptr<GaussiansBuffer>> models, int queues, uint64_t hbm_budget)
    : dev_(dev), models_(std::move(models)), budget_(hbm_budget) {
  if (queues < 1) queues = 1;
  queues_.push_back(0);
  for (int q = 1; q < queues; ++q) queues_.push_back(dev.add_queue());
}

Server::~Server() {
  try {
    dev_.wait();
  } catch (...) {
  }
  for (auto& s : slots_)
    if (s->done) hipEventDestroy(s->done);
}

uint64_t Server::footprint(uint32_t model, uint32_t w, uint32_t h) const {
  // per-view buffers of a Viewer (indices, keys, splats, sort and tile-list
  // scratch) plus the target; measured sizes replace this once allocated
  const uint64_t n = models_[model]->len();
  return n * 120u + uint64_t(w) * h * 4u + (64ull << 20);
}

Server::Slot* Server::acquire(const ServeRequest& r) {
  for (auto& s : slots_)
    if (!s->busy && s->model == r.model && s->width == r.width && s->height == r.height) return s.get();
  uint64_t need = footprint(r.model, r.width, r.height);
  // make room by dropping idle slots of other shapes
  while (used_ + need > budget_) {
    auto it = std::find_if(slots_.begin(), slots_.end(), [](const std::unique_ptr<Slot>& s) { return !s->busy; });
    if (it == slots_.end()) return nullptr;
    used_ -= (*it)->bytes;
    hipEventDestroy((*it)->done);
    slots_.erase(it);
  }
  auto s = std::make_unique<Slot>();
  s->model = r.model;
  s->width = r.width;
  s->height = r.height;
  const int q = queues_[slots_.size() % queues_.size()];
  const uint64_t before = dev_.allocated();
  s->viewer = std::make_unique<Viewer>(dev_, kFormatRgba8Unorm, models_[r.model], q);
  s->target = std::make_unique<Texture>(dev_, r.width, r.height, kFormatRgba8Unorm);
  s->bytes = std::max<uint64_t>(dev_.allocated() - before, 1);
  hip_check(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "hipEventCreate");
  used_ += s->bytes;
  slots_.push_back(std::move(s));
  return slots_.back().get();
}

void Server::release(Slot* s) { s->busy = false; }

ServeStats Server::run(const std::vector<ServeRequest>& reqs, const std::vector<uint8_t*>* images) {
  using clock = std::chrono::steady_clock;
  ServeStats st;
  st.requests = reqs.size();
  std::vector<double> lat(reqs.size(), 0.0);
  std::deque<size_t> pending;
  std::vector<Slot*> in_flight;
  size_t next = 0;
  const size_t max_in_flight = 2 * queues_.size();
  double in_flight_area = 0.0, last_t = 0.0;
  const auto t0 = clock::now();
  auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
  double first_arrival = reqs.empty() ? 0.0 : reqs.front().arrival_s, last_done = 0.0;

  while (st.completed < reqs.size()) {
    const double now = now_s();
    in_flight_area += double(in_flight.size()) * (now - last_t);
    last_t = now;
    while (next < reqs.size() && reqs[next].arrival_s <= now) pending.push_back(next++);
    bool progressed = false;
    // completions
    for (size_t k = 0; k < in_flight.size();) {
      Slot* s = in_flight[k];
      const hipError_t q = hipEventQuery(s->done);
      if (q == hipErrorNotReady) {
        ++k;
        continue;
      }
      hip_check(q, "hipEventQuery");
      hipStream_t sq = dev_.stream(s->viewer->queue);
      // a frame whose tile lists overflowed is drawn again with larger lists
      s->viewer->renderer.settle();
      const double t = now_s();
      lat[s->request] = t - reqs[s->request].arrival_s;
      last_done = std::max(last_done, t);
      if (images && (*images)[s->request]) {
        hip_check(hipMemcpyAsync((*images)[s->request], s->target->dev(), s->target->bytes(), hipMemcpyDeviceToHost, sq),
                  "hipMemcpyAsync(image)");
        hip_check(hipStreamSynchronize(sq), "hipStreamSynchronize");
      }
      release(s);
      in_flight.erase(in_flight.begin() + long(k));
      ++st.completed;
      progressed = true;
    }
    // admissions, first come first served
    while (!pending.empty() && in_flight.size() < max_in_flight) {
      const ServeRequest& r = reqs[pending.front()];
      Slot* s = acquire(r);
      if (!s) {
        // with nothing in flight no completion will free memory for it
        if (in_flight.empty())
          throw Error(GS3D_ERROR_OUT_OF_MEMORY, "a request does not fit the server's HBM budget");
        ++st.deferred_for_memory;
        break;
      }
      s->busy = true;
      s->request = pending.front();
      pending.pop_front();
      s->viewer->update_camera_with_pod(r.pod);
      s->viewer->render(*s->target, false);
      hip_check(hipEventRecord(s->done, dev_.stream(s->viewer->queue)), "hipEventRecord");
      in_flight.push_back(s);
      st.peak_hbm_bytes = std::max<uint64_t>(st.peak_hbm_bytes, used_);
      progressed = true;
    }
    if (!progressed) {
      if (in_flight.empty() && pending.empty() && next < reqs.size()) {
        const double wait = reqs[next].arrival_s - now_s();
        if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
      } else {
        std::this_thread::yield();
      }
    }
  }
  const uint32_t e = dev_.take_errors();
  throw_kernel_errors(e);
  st.seconds = std::max(last_done - first_arrival, 1e-9);
  st.frames_per_s = double(st.completed) / st.seconds;
  st.mean_in_flight = in_flight_area / std::max(last_t, 1e-9);
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  if (!sorted.empty()) {
    auto pct = [&](double p) {
      const double idx = p * double(sorted.size() - 1);
      const size_t lo = size_t(idx);
      const size_t hi = std::min(lo + 1, sorted.size() - 1);
      return 1e3 * (sorted[lo] + (sorted[hi] - sorted[lo]) * (idx - double(lo)));
    };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    st.max_ms = 1e3 * sorted.back();
    double sum = 0.0;
    for (double x : sorted) sum += x;
    st.mean_ms = 1e3 * sum / double(sorted.size());
  }
  return st;
}

void Server::render_alone(const ServeRequest& r, uint8_t* out) {
  dev_.wait();
  Slot* s = acquire(r);
  if (!s) throw Error(GS3D_ERROR_OUT_OF_MEMORY, "the request does not fit the HBM budget");
  s->viewer->update_camera_with_pod(r.pod);
  s->viewer->render(*s->target, false);
  hip_check(hipStreamSynchronize(dev_.stream(s->viewer->queue)), "hipStreamSynchronize");
  s->viewer->renderer.settle();
  hip_check(hipMemcpy(out, s->target->dev(), s->target->bytes(), hipMemcpyDeviceToHost), "hipMemcpy(image)");
}

}  // namespace gs3d
