// isp_server.hip - the served form (include/isp_c.h, isp_server_*): continuous
// batching of many cameras' frames on one GPU.
//
// A scheduling thread owns every HIP call of the server. Each round it
//  1. retires launches whose event has completed (their frames become results);
//  2. for each free launch slot, admits waiting requests in arrival order: a
//     request is taken when its camera has no frame in flight and the arena can
//     hold its buffers; the first request that does not fit stops admission for
//     this round, so large frames are not starved by smaller ones behind them;
//  3. copies the admitted host frames into the arena on the slot's stream and
//     launches them as one batch (a lone frame, or every frame in the baseline
//     mode, as a per-frame launch).
// Launch slots are independent streams, so a launch starts while earlier ones
// still run. A camera is one pipeline (tables, white balance and exposure
// state on the device); at most one of its frames is in flight, so its frames
// run in order.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "isp_pipeline.hpp"

using namespace isp_amd;

namespace {

constexpr std::size_t kAlign = 256;

std::size_t round_up(std::size_t v, std::size_t a) { return (v + a - 1) / a * a; }

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// First-fit allocator over one device allocation: the admission budget.
class Arena {
 public:
  void init(std::size_t bytes) {
    size_ = bytes;
    free_.clear();
    if (bytes > 0) free_[0] = bytes;
  }
  // Offset of `bytes` (rounded to kAlign), or -1 if no free range holds it.
  long long take(std::size_t bytes) {
    bytes = round_up(bytes, kAlign);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const std::size_t at = it->first;
      const std::size_t left = it->second - bytes;
      free_.erase(it);
      if (left > 0) free_[at + bytes] = left;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return static_cast<long long>(at);
    }
    return -1;
  }
  void give(std::size_t at, std::size_t bytes) {
    bytes = round_up(bytes, kAlign);
    used_ -= bytes;
    auto next = free_.lower_bound(at);
    if (next != free_.end() && at + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == at) {
        prev->second += bytes;
        return;
      }
    }
    free_[at] = bytes;
  }
  std::size_t used() const { return used_; }
  std::size_t peak() const { return peak_; }
  std::size_t size() const { return size_; }

 private:
  std::size_t size_ = 0;
  std::size_t used_ = 0;
  std::size_t peak_ = 0;
  std::map<std::size_t, std::size_t> free_;
};

struct Camera {
  isp_camera_config config{};
  isp_pipeline* pipeline = nullptr;
  std::size_t raw_pitch = 0;  // of the arena copy
  std::size_t raw_bytes = 0;
  std::size_t out_pitch = 0;
  std::size_t out_bytes = 0;
  std::size_t stats_bytes = 0;  // histograms then the statistics vector
  bool busy = false;            // a frame of this camera is in flight
};

struct Request {
  int64_t id = 0;
  int32_t camera = 0;
  uint64_t tag = 0;
  const void* raw = nullptr;
  std::size_t pitch = 0;
  uint8_t location = ISP_LOCATION_HOST;
  long long at = -1;  // arena offset, once admitted
  std::size_t bytes = 0;
  isp_server_result result{};
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  isp_batch* batch = nullptr;
  std::vector<Request*> frames;  // in flight
  bool busy = false;
};

}  // namespace

struct isp_server {
  isp_server_config config{};
  int device = 0;
  std::uint8_t* arena_base = nullptr;
  Arena arena;
  std::vector<Camera> cameras;
  std::vector<Slot> slots;
  std::deque<Request*> waiting;
  std::deque<Request*> finished;
  std::map<int64_t, Request*> held;  // finished, polled, not yet released
  int64_t next_id = 1;
  isp_server_counters counters{};
  int in_flight = 0;
  unsigned long long epoch = 0;  // advances on every submit and release
  std::mutex mutex;
  std::condition_variable wake;      // the scheduling thread
  std::condition_variable finished_cv;  // pollers
  std::thread worker;
  bool stop = false;
  std::string failure;  // a HIP failure of the scheduling thread

  void run() {
    check(hipSetDevice(device), "hipSetDevice");
    std::unique_lock<std::mutex> lock(mutex);
    unsigned long long seen = ~0ULL;  // the epoch of the last admission scan
    while (true) {
      bool progress = retire();
      if (progress || seen != epoch) {
        seen = epoch;
        for (Slot& slot : slots) {
          if (!slot.busy && !waiting.empty()) progress = launch(slot) || progress;
        }
      }
      if (stop && in_flight == 0 && (waiting.empty() || !progress)) break;
      if (progress) continue;
      if (in_flight > 0) {
        // Launches on the GPU: poll their events closely.
        lock.unlock();
        std::this_thread::yield();
        lock.lock();
      } else {
        // Nothing on the GPU and nothing admissible: sleep until a submit or a release.
        wake.wait(lock, [&] { return stop || seen != epoch; });
      }
    }
    // Requests that can no longer be admitted at shutdown fail.
    for (Request* r : waiting) {
      r->result.status = ISP_FAILURE;
      r->result.done_ns = now_ns();
      finished.push_back(r);
    }
    waiting.clear();
  }

  // Turns completed launches into results.
  bool retire() {
    bool any = false;
    for (Slot& slot : slots) {
      if (!slot.busy) continue;
      const hipError_t q = hipEventQuery(slot.done);
      if (q == hipErrorNotReady) continue;
      const int64_t t = now_ns();
      std::string error;
      if (q != hipSuccess) error = std::string("IspPipeline (hip): launch: ") + hipGetErrorString(q);
      for (Request* r : slot.frames) {
        r->result.done_ns = t;
        if (!error.empty()) r->result.status = ISP_FAILURE;
        cameras[r->camera].busy = false;
        finished.push_back(r);
        ++counters.completed;
      }
      if (!error.empty()) failure = error;
      slot.frames.clear();
      slot.busy = false;
      --in_flight;
      any = true;
    }
    if (any) finished_cv.notify_all();
    return any;
  }

  // Admits what fits into this slot's launch and launches it.
  bool launch(Slot& slot) {
    std::vector<Request*> take;
    for (auto it = waiting.begin(); it != waiting.end() && static_cast<int>(take.size()) < config.max_batch;) {
      Request* r = *it;
      Camera& cam = cameras[r->camera];
      if (cam.busy) {
        ++it;
        continue;
      }
      const std::size_t bytes =
          (r->location == ISP_LOCATION_HOST ? round_up(cam.raw_bytes, kAlign) : 0) + round_up(cam.out_bytes, kAlign) +
          round_up(cam.stats_bytes, kAlign);
      const long long at = arena.take(bytes);
      if (at < 0) break;  // first come, first served: nothing behind it jumps ahead
      r->at = at;
      r->bytes = bytes;
      cam.busy = true;
      take.push_back(r);
      it = waiting.erase(it);
    }
    if (take.empty()) return false;
    counters.budget_used = arena.used();
    counters.budget_peak = arena.peak();
    counters.queued = static_cast<int64_t>(waiting.size());

    // The HIP work of the launch (views into the arena).
    const int64_t t = now_ns();
    std::vector<isp_batch_frame> frames(take.size());
    for (std::size_t i = 0; i < take.size(); ++i) {
      Request* r = take[i];
      Camera& cam = cameras[r->camera];
      std::uint8_t* p = arena_base + r->at;
      const void* raw = r->raw;
      std::size_t raw_pitch = r->pitch;
      if (r->location == ISP_LOCATION_HOST) {
        const std::size_t row = raw_row_bytes(cam.config.format, cam.config.width);
        check(hipMemcpy2DAsync(p, cam.raw_pitch, r->raw, r->pitch, row, cam.config.height, hipMemcpyHostToDevice,
                               slot.stream),
              "hipMemcpy2DAsync");
        raw = p;
        raw_pitch = cam.raw_pitch;
        p += round_up(cam.raw_bytes, kAlign);
      }
      isp_batch_frame& f = frames[i];
      f.pipeline = cam.pipeline;
      f.raw = isp_raw_view{raw, cam.config.width, cam.config.height, raw_pitch, cam.config.format,
                           cam.config.bit_depth, cam.config.pattern};
      f.out = isp_rgb_view{p, cam.out_pitch, cam.config.out_format};
      std::uint8_t* s = p + round_up(cam.out_bytes, kAlign);
      f.stats = cam.config.stats ? isp_stats_out{reinterpret_cast<uint32_t*>(s),
                                                 reinterpret_cast<float*>(s + cam.stats_bytes - ISP_STATS_SIZE * 4)}
                                 : isp_stats_out{nullptr, nullptr};
      f.exposure = cam.config.initial;
      r->result.out = f.out.data;
      r->result.out_pitch = cam.out_pitch;
      r->result.histogram = f.stats.histogram;
      r->result.stats = f.stats.stats;
      r->result.launch_ns = t;
      r->result.batch_frames = static_cast<int32_t>(take.size());
    }
    std::string error;
    // A launch of one frame runs the per-frame kernels: for a lone frame they
    // start faster than the batch kernels' work queue.
    if (config.per_frame || frames.size() == 1) {
      for (const isp_batch_frame& f : frames) {
        if (isp_process(f.pipeline, &f.raw, &f.out, &f.stats, &f.exposure, slot.stream) != ISP_OK) {
          error = isp_last_error();
          break;
        }
      }
    } else if (isp_batch_set_frames(slot.batch, frames.data(), static_cast<int32_t>(frames.size())) != ISP_OK ||
               isp_batch_launch(slot.batch, slot.stream) != ISP_OK) {
      error = isp_last_error();
    }
    check(hipEventRecord(slot.done, slot.stream), "hipEventRecord");
    slot.frames = take;
    slot.busy = true;
    ++in_flight;
    ++counters.launches;
    counters.max_in_flight = std::max(counters.max_in_flight, in_flight);
    if (!error.empty()) {
      // The frames fail; they still go through retire() so their memory returns.
      for (Request* r : take) r->result.status = ISP_FAILURE;
      failure = error;
    }
    return true;
  }

  // Admission in the submitting thread (the mutex held): while a launch slot is
  // free, waiting frames go to the GPU now, not when the scheduling thread next
  // wakes up (a sleeping thread's wake-up takes tens of microseconds; at low
  // load a frame's whole latency is under a hundred). The scheduling thread
  // still retires the launches. A HIP failure here stops the server, as one in
  // the scheduling thread does.
  void admit_here() {
    if (stop || !failure.empty()) return;
    int previous = -1;
    if (hipGetDevice(&previous) != hipSuccess) return;  // left to the scheduling thread
    if (previous != device && hipSetDevice(device) != hipSuccess) return;
    try {
      for (Slot& slot : slots) {
        if (!slot.busy && !waiting.empty()) launch(slot);
      }
    } catch (const std::exception& e) {
      failure = e.what();
      stop = true;
    }
    if (previous != device) static_cast<void>(hipSetDevice(previous));
  }

  void shutdown() noexcept {
    {
      std::lock_guard<std::mutex> g(mutex);
      stop = true;
    }
    wake.notify_all();
    if (worker.joinable()) worker.join();
    for (Slot& slot : slots) {
      if (slot.batch != nullptr) static_cast<void>(isp_batch_destroy(slot.batch));
      if (slot.done != nullptr) static_cast<void>(hipEventDestroy(slot.done));
      if (slot.stream != nullptr) static_cast<void>(hipStreamDestroy(slot.stream));
    }
    slots.clear();
    for (Camera& cam : cameras) static_cast<void>(isp_destroy(cam.pipeline));
    cameras.clear();
    for (Request* r : finished) delete r;
    for (auto& [id, r] : held) delete r;
    for (Request* r : waiting) delete r;
    finished.clear();
    held.clear();
    waiting.clear();
    if (arena_base != nullptr) static_cast<void>(hipFree(arena_base));
    arena_base = nullptr;
  }
};

extern "C" {

int64_t isp_server_now_ns(void) { return now_ns(); }

void isp_server_config_init(isp_server_config* config) {
  if (config == nullptr) return;
  std::memset(config, 0, sizeof(*config));
  config->hbm_budget = 256ULL << 20;
  config->max_batch = 32;
  config->launch_slots = 2;
  config->per_frame = 0;
  isp_batch_schedule_init(&config->schedule);
  config->schedule.graph = 0;  // every launch has a different frame set
  config->schedule.queue = ISP_BATCH_QUEUE_STATIC;  // small launches: no queue atomics
}

isp_status isp_server_create(const isp_server_config* config, isp_server** out) {
  return guarded([&] {
    if (config == nullptr || out == nullptr) throw invalid_argument("config or out is null");
    *out = nullptr;
    if (config->hbm_budget < kAlign) throw invalid_argument("hbm_budget must be at least 256 bytes");
    if (config->max_batch < 1 || config->max_batch > ISP_BATCH_MAX_FRAMES) {
      throw invalid_argument("max_batch must be in 1.." + std::to_string(ISP_BATCH_MAX_FRAMES));
    }
    if (config->launch_slots < 1 || config->launch_slots > 8) throw invalid_argument("launch_slots must be in 1..8");
    if (config->per_frame != 0 && config->per_frame != 1) throw invalid_argument("per_frame must be 0 or 1");
    auto s = std::make_unique<isp_server>();
    s->config = *config;
    check(hipGetDevice(&s->device), "hipGetDevice");
    check(hipMalloc(&s->arena_base, config->hbm_budget), "hipMalloc (the serving budget)");
    s->arena.init(config->hbm_budget);
    s->slots.resize(config->launch_slots);
    for (Slot& slot : s->slots) {
      check(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking), "hipStreamCreate");
      check(hipEventCreateWithFlags(&slot.done, kOrderingEventFlags), "hipEventCreate");
      if (config->per_frame == 0) {
        if (isp_batch_create(&slot.batch) != ISP_OK) throw runtime_failure(isp_last_error());
        if (isp_batch_set_schedule(slot.batch, &config->schedule) != ISP_OK) throw invalid_argument(isp_last_error());
      }
    }
    isp_server* raw = s.get();
    s->worker = std::thread([raw] {
      try {
        raw->run();
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(raw->mutex);
        raw->failure = e.what();
        raw->stop = true;
      }
      raw->finished_cv.notify_all();
    });
    *out = s.release();
  });
}

isp_status isp_server_destroy(isp_server* server) {
  return guarded([&] {
    if (server == nullptr) return;
    server->shutdown();
    delete server;
  });
}

isp_status isp_server_add_camera(isp_server* server, const isp_camera_config* camera, int32_t* id) {
  return guarded([&] {
    if (server == nullptr || camera == nullptr || id == nullptr) throw invalid_argument("server, camera or id is null");
    Camera cam;
    cam.config = *camera;
    const isp_raw_view probe{nullptr, camera->width, camera->height, raw_row_bytes(camera->format, camera->width),
                             camera->format, camera->bit_depth, camera->pattern};
    Params p = copy_params(&camera->params);
    {
      isp_raw_view v = probe;
      v.data = reinterpret_cast<const void*>(16);  // validation reads no memory
      validate_raw(v, p.c);
    }
    if (camera->out_format > ISP_OUT_RGB32F) throw invalid_argument("out_format is not an OutputFormat");
    const std::size_t row = raw_row_bytes(camera->format, camera->width);
    cam.raw_pitch = round_up(row, 16);
    cam.raw_bytes = cam.raw_pitch * static_cast<std::size_t>(camera->height);
    const std::size_t out_row = static_cast<std::size_t>(camera->width) * isp_output_channels(camera->out_format) *
                                isp_output_channel_bytes(camera->out_format);
    cam.out_pitch = round_up(out_row, 16);
    cam.out_bytes = cam.out_pitch * static_cast<std::size_t>(camera->height);
    cam.stats_bytes = camera->stats ? round_up(ISP_HISTOGRAM_CHANNELS * 4ULL * p.bins(), 16) + ISP_STATS_SIZE * 4 : 0;
    {
      const isp_rgb_view o{reinterpret_cast<void*>(16), cam.out_pitch, camera->out_format};
      isp_raw_view v = probe;
      v.data = reinterpret_cast<void*>(16);
      v.pitch = cam.raw_pitch;
      validate_out(v, o, p.c);
    }
    const std::size_t need = round_up(cam.raw_bytes, kAlign) + round_up(cam.out_bytes, kAlign) +
                             round_up(cam.stats_bytes, kAlign);
    if (need > server->config.hbm_budget) {
      throw out_of_budget("camera frames need " + std::to_string(need) + " bytes, more than the serving budget");
    }
    std::lock_guard<std::mutex> g(server->mutex);
    if (isp_create(&camera->params, ISP_BACKEND_GPU, &cam.pipeline) != ISP_OK) {
      throw invalid_argument(isp_last_error());
    }
    if (camera->initial.exposure_ns > 0 &&
        isp_set_exposure_feedback(cam.pipeline, 1, &camera->initial) != ISP_OK) {
      const std::string e = isp_last_error();
      isp_destroy(cam.pipeline);
      throw invalid_argument(e);
    }
    server->cameras.push_back(cam);
    *id = static_cast<int32_t>(server->cameras.size() - 1);
  });
}

isp_status isp_server_submit(isp_server* server, int32_t camera, const void* raw, size_t pitch, uint8_t location,
                             uint64_t tag, int64_t* request) {
  return guarded([&] {
    if (server == nullptr || raw == nullptr) throw invalid_argument("server or raw is null");
    if (location != ISP_LOCATION_HOST && location != ISP_LOCATION_DEVICE) {
      throw invalid_argument("location is not a MemoryLocation");
    }
    auto r = std::make_unique<Request>();
    {
      std::lock_guard<std::mutex> g(server->mutex);
      if (!server->failure.empty()) throw runtime_failure(server->failure);
      if (camera < 0 || camera >= static_cast<int32_t>(server->cameras.size())) {
        throw invalid_argument("camera is not a camera of this server");
      }
      const Camera& cam = server->cameras[camera];
      if (pitch < raw_row_bytes(cam.config.format, cam.config.width)) {
        throw invalid_argument("raw pitch is shorter than a row");
      }
      r->id = server->next_id++;
      r->camera = camera;
      r->tag = tag;
      r->raw = raw;
      r->pitch = pitch;
      r->location = location;
      r->result.request = r->id;
      r->result.camera = camera;
      r->result.tag = tag;
      r->result.submit_ns = now_ns();
      if (request != nullptr) *request = r->id;
      server->waiting.push_back(r.release());
      ++server->epoch;
      ++server->counters.submitted;
      server->counters.queued = static_cast<int64_t>(server->waiting.size());
      server->admit_here();
    }
    server->wake.notify_one();
  });
}

isp_status isp_server_poll(isp_server* server, int64_t timeout_us, isp_server_result* result) {
  if (server == nullptr || result == nullptr) {
    set_error("server or result is null");
    return ISP_INVALID_ARGUMENT;
  }
  std::unique_lock<std::mutex> lock(server->mutex);
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::microseconds(std::max<int64_t>(0, timeout_us));
  server->finished_cv.wait_until(lock, deadline, [&] { return !server->finished.empty() || server->stop; });
  if (server->finished.empty()) {
    if (!server->failure.empty()) {
      set_error(server->failure);
      return ISP_FAILURE;
    }
    set_error("no finished frame yet");
    return ISP_NOT_READY;
  }
  Request* r = server->finished.front();
  server->finished.pop_front();
  server->held[r->id] = r;
  *result = r->result;
  clear_error();
  return ISP_OK;
}

isp_status isp_server_release(isp_server* server, int64_t request) {
  return guarded([&] {
    if (server == nullptr) throw invalid_argument("server is null");
    {
      std::lock_guard<std::mutex> g(server->mutex);
      auto it = server->held.find(request);
      if (it == server->held.end()) throw invalid_argument("request is not a polled, unreleased result");
      Request* r = it->second;
      server->arena.give(static_cast<std::size_t>(r->at), r->bytes);
      server->counters.budget_used = server->arena.used();
      server->held.erase(it);
      delete r;
      ++server->epoch;
    }
    server->wake.notify_one();
  });
}

isp_status isp_server_counters_get(isp_server* server, isp_server_counters* counters) {
  return guarded([&] {
    if (server == nullptr || counters == nullptr) throw invalid_argument("server or counters is null");
    std::lock_guard<std::mutex> g(server->mutex);
    *counters = server->counters;
    counters->queued = static_cast<int64_t>(server->waiting.size());
    counters->budget_used = server->arena.used();
    counters->budget_peak = server->arena.peak();
  });
}

}  // extern "C"
