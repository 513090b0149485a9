// isp_batch.hip - the batched form (include/isp_c.h, isp_batch_*): many camera
// streams, each a pipeline with its own tables and state, in two persistent
// launches (white balance statistics of every auto-balanced frame, then every
// frame's main pass; isp_kernels.hip), optionally replayed as a hipGraph, with
// an autotuner over the launch schedule.
//
// The host builds the work queue's table (one BatchFrame per frame plus the
// order in which the queue lists them) in pinned memory and copies it to the
// device ahead of the launch that first uses it. The kernel's control block
// resets itself at the end of every launch, so a captured launch replays with
// no memset node.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <numeric>
#include <sstream>

#include "isp_pipeline.hpp"

using namespace isp_amd;

namespace {

constexpr int kDefaultTilesPerItem = 4;
constexpr int kDefaultAwbTilesPerItem = 8;

void schedule_defaults(isp_batch_schedule* s) {
  std::memset(s, 0, sizeof(*s));
  for (auto& row : s->tile_height) row[0] = row[1] = 32;
  for (auto& row : s->tiles_per_item) row[0] = row[1] = kDefaultTilesPerItem;
  s->awb_tiles_per_item = kDefaultAwbTilesPerItem;
  s->blocks_per_cu = 0;
  s->order = ISP_BATCH_ORDER_LARGEST_FIRST;
  s->queue = ISP_BATCH_QUEUE_DYNAMIC;
  s->graph = 1;
  s->per_frame = 0;
}

void validate_schedule(const isp_batch_schedule& s) {
  for (const auto& row : s.tile_height) {
    for (int h : row) {
      if (h != 16 && h != 32) throw invalid_argument("schedule.tile_height must be 16 or 32");
    }
  }
  for (const auto& row : s.tiles_per_item) {
    for (int n : row) {
      if (n < 1 || n > 32) throw invalid_argument("schedule.tiles_per_item must be in 1..32");
    }
  }
  if (s.awb_tiles_per_item < 1 || s.awb_tiles_per_item > 64) {
    throw invalid_argument("schedule.awb_tiles_per_item must be in 1..64");
  }
  if (s.blocks_per_cu < 0 || s.blocks_per_cu > 8) throw invalid_argument("schedule.blocks_per_cu must be in 0..8");
  if (s.order != ISP_BATCH_ORDER_LARGEST_FIRST && s.order != ISP_BATCH_ORDER_AS_GIVEN) {
    throw invalid_argument("schedule.order is not an ISP_BATCH_ORDER");
  }
  if (s.queue < ISP_BATCH_QUEUE_DYNAMIC || s.queue > ISP_BATCH_QUEUE_STATIC_XCD) {
    throw invalid_argument("schedule.queue is not an ISP_BATCH_QUEUE");
  }
  if (s.graph != 0 && s.graph != 1) throw invalid_argument("schedule.graph must be 0 or 1");
  if (s.per_frame != 0 && s.per_frame != 1) throw invalid_argument("schedule.per_frame must be 0 or 1");
}

bool aligned16(const void* p, long long pitch) {
  return ((reinterpret_cast<unsigned long long>(p) | static_cast<unsigned long long>(pitch)) & 15ULL) == 0;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace

struct isp_batch {
  int device = 0;
  int cu_count = 0;
  isp_batch_schedule schedule{};
  std::vector<isp_batch_frame> frames;
  BatchTable* h_table = nullptr;  // pinned
  BatchTable* d_table = nullptr;
  BatchControl* d_control = nullptr;
  std::size_t table_bytes = 0;
  bool dirty = false;    // the host table is newer than the device's
  bool sharpen = false;  // a frame sharpens: the batch kernel with the larger LDS
  bool wide = true;      // every frame runs 32-row tiles: the 512-thread kernels
  int uniform = -1;      // every frame runs this variant with 32-row tiles: its kernel alone
  // Frames whose tone table starts at -0.0 (FrameDesc::zero_sign): they run the
  // per-frame kernels, which have the form that keeps every zero's sign, in the
  // same launch sequence, ahead of the batch kernels; the batch kernels carry
  // only the common form.
  std::vector<int> solo;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_grid = 0;  // the grids the graph was captured with
  hipEvent_t done = nullptr;
  bool queued = false;

  ~isp_batch() { release(); }

  void release() noexcept {
    if (queued && done != nullptr) static_cast<void>(hipEventSynchronize(done));
    drop_graph();
    if (h_table != nullptr) static_cast<void>(hipHostFree(h_table));
    if (d_table != nullptr) static_cast<void>(hipFree(d_table));
    if (d_control != nullptr) static_cast<void>(hipFree(d_control));
    if (capture != nullptr) static_cast<void>(hipStreamDestroy(capture));
    if (done != nullptr) static_cast<void>(hipEventDestroy(done));
    h_table = nullptr;
    d_table = nullptr;
    d_control = nullptr;
    capture = nullptr;
    done = nullptr;
    queued = false;
  }

  void drop_graph() noexcept {
    if (exec != nullptr) static_cast<void>(hipGraphExecDestroy(exec));
    if (graph != nullptr) static_cast<void>(hipGraphDestroy(graph));
    exec = nullptr;
    graph = nullptr;
    graph_grid = 0;
  }

  void wait() {
    if (queued) check(hipEventSynchronize(done), "hipEventSynchronize");
  }

  int resident_per_cu() const {
    const int fit = batch_main_blocks_per_cu(sharpen, wide, uniform);
    return schedule.blocks_per_cu > 0 ? std::min(schedule.blocks_per_cu, fit) : fit;
  }

  int main_grid() const { return std::max(1, std::min(h_table->main_items, cu_count * resident_per_cu())); }
  int awb_grid() const { return std::max(1, std::min(h_table->awb_items, cu_count * batch_awb_blocks_per_cu())); }

  // Builds the host table for `frames` under `schedule`.
  void build() {
    BatchTable& t = *h_table;
    const int n = static_cast<int>(frames.size());
    std::memset(&t, 0, sizeof(BatchTable));
    t.count = n;
    t.queue = schedule.queue;
    sharpen = false;
    wide = true;
    uniform = -1;
    int first_variant = -1;
    bool one_variant = true;
    solo.clear();
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
      const isp_batch_frame& fr = frames[i];
      const isp_pipeline* p = fr.pipeline;
      const isp_params& c = p->params.c;
      BatchFrame& b = t.frames[i];
      b.f = p->frame(fr.raw, fr.out, &fr.stats, &fr.exposure);
      b.s = p->d_scratch;
      const bool malvar = c.demosaic == ISP_DEMOSAIC_MALVAR;
      const bool sharpen = c.sharpen > 0.0F;
      const int tile_h = schedule.tile_height[fr.raw.format][malvar ? 1 : 0];
      if (b.f.zero_sign != 0 || schedule.per_frame != 0) {
        // No work items in the batch passes: launch() runs the frame by itself.
        solo.push_back(i);
        cost[i] = 0.0;
        continue;
      }
      this->sharpen = this->sharpen || sharpen;
      wide = wide && tile_h == 32;
      b.variant = fr.raw.format | (malvar ? 4 : 0) | (sharpen ? 8 : 0) | (tile_h == 16 ? 16 : 0);
      if (first_variant < 0) first_variant = b.variant;
      one_variant = one_variant && b.variant == first_variant;
      b.tiles_x = ceil_div(fr.raw.width, kBatchTileW);
      b.tiles = b.tiles_x * ceil_div(fr.raw.height, tile_h);
      b.tiles_per_item = schedule.tiles_per_item[fr.raw.format][malvar ? 1 : 0];
      b.items = ceil_div(b.tiles, b.tiles_per_item);
      if (c.white_balance != ISP_WB_MANUAL) {
        b.awb_tiles_x = ceil_div(fr.raw.width, kBatchAwbTile);
        b.awb_tiles = b.awb_tiles_x * ceil_div(fr.raw.height, kBatchAwbTile);
        b.awb_tiles_per_item = schedule.awb_tiles_per_item;
        b.awb_items = ceil_div(b.awb_tiles, b.awb_tiles_per_item);
      }
      b.fast_rows = aligned16(fr.raw.data, static_cast<long long>(fr.raw.pitch)) ? 1 : 0;
      b.out_aligned = aligned16(fr.out.data, static_cast<long long>(fr.out.pitch)) ? 1 : 0;
      cost[i] = static_cast<double>(fr.raw.width) * fr.raw.height * (sharpen ? 1.6 : 1.0) * (malvar ? 1.2 : 1.0);
    }
    if (wide && one_variant && first_variant >= 0) uniform = first_variant;
    // A small launch gets smaller work items, so that every resident block has
    // work: the items per pass are kept at twice the blocks or more when the
    // tiles allow it, never above the schedule's item size.
    const int capacity = cu_count * resident_per_cu();
    const int awb_capacity = cu_count * batch_awb_blocks_per_cu();
    long long main_tiles = 0;
    long long awb_tiles = 0;
    for (int i = 0; i < n; ++i) {
      main_tiles += t.frames[i].tiles;
      awb_tiles += t.frames[i].awb_tiles;
    }
    const int main_cap = static_cast<int>(std::max(1LL, main_tiles / (2LL * capacity)));
    const int awb_cap = static_cast<int>(std::max(1LL, awb_tiles / (2LL * awb_capacity)));
    for (int i = 0; i < n; ++i) {
      BatchFrame& b = t.frames[i];
      if (b.tiles == 0) continue;  // a frame launch() runs by itself
      b.tiles_per_item = std::min(b.tiles_per_item, main_cap);
      b.items = ceil_div(b.tiles, b.tiles_per_item);
      if (b.awb_tiles > 0) {
        b.awb_tiles_per_item = std::min(b.awb_tiles_per_item, awb_cap);
        b.awb_items = ceil_div(b.awb_tiles, b.awb_tiles_per_item);
      }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (schedule.order == ISP_BATCH_ORDER_LARGEST_FIRST) {
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    }
    int awb = 0;
    int main = 0;
    for (int k = 0; k < n; ++k) {
      t.awb_order[k] = order[k];
      t.awb_start[k] = awb;
      awb += t.frames[order[k]].awb_items;
      t.main_order[k] = order[k];
      t.main_start[k] = main;
      main += t.frames[order[k]].items;
    }
    t.awb_items = awb;
    t.main_items = main;
    table_bytes = offsetof(BatchTable, frames) + static_cast<std::size_t>(n) * sizeof(BatchFrame);
    dirty = true;
    drop_graph();  // the grid may change with the items
  }

  void launch(hipStream_t stream) {
    static_cast<void>(hipGetLastError());
    if (frames.empty()) throw invalid_argument("the batch has no frames (isp_batch_set_frames)");
    for (const isp_batch_frame& fr : frames) {
      if (fr.pipeline->queued) check(hipStreamWaitEvent(stream, fr.pipeline->done, 0), "hipStreamWaitEvent");
    }
    if (queued) check(hipStreamWaitEvent(stream, done, 0), "hipStreamWaitEvent");
    if (dirty) {
      check(hipMemcpyAsync(d_table, h_table, table_bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
      dirty = false;
    }
    const int awb_blocks = h_table->awb_items > 0 ? awb_grid() : 0;
    const int main_blocks = main_grid();
    const auto enqueue = [&](hipStream_t st) {
      for (int i : solo) {
        const BatchFrame& b = h_table->frames[i];
        const isp_pipeline* p = frames[i].pipeline;
        if (p->params.c.white_balance != ISP_WB_MANUAL) launch_awb(b.f, b.s, p->launch, st);
        launch_main(b.f, b.s, p->params.c.demosaic == ISP_DEMOSAIC_MALVAR, p->params.c.sharpen > 0.0F, p->launch, st);
      }
      if (awb_blocks > 0) launch_batch_awb(d_table, d_control, awb_blocks, st);
      if (h_table->main_items > 0) launch_batch_main(d_table, d_control, main_blocks, sharpen, wide, uniform, st);
    };
    if (schedule.graph != 0) {
      if (exec == nullptr || graph_grid != main_blocks * 65536 + awb_blocks) {
        drop_graph();
        check(hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
          enqueue(capture);
        } catch (...) {
          hipGraph_t partial = nullptr;
          static_cast<void>(hipStreamEndCapture(capture, &partial));
          if (partial != nullptr) static_cast<void>(hipGraphDestroy(partial));
          throw;
        }
        check(hipStreamEndCapture(capture, &graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        graph_grid = main_blocks * 65536 + awb_blocks;
      }
      check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    } else {
      enqueue(stream);
    }
    check(hipEventRecord(done, stream), "hipEventRecord");
    queued = true;
    for (const isp_batch_frame& fr : frames) {
      check(hipEventRecord(fr.pipeline->done, stream), "hipEventRecord");
      fr.pipeline->queued = true;
      fr.pipeline->last_stream = stream;
    }
  }

  // The autotuner's cache key: the device and the frame set, order-independent.
  std::string key() const {
    hipDeviceProp_t prop{};
    check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    std::vector<std::string> parts;
    for (const isp_batch_frame& fr : frames) {
      const isp_params& c = fr.pipeline->params.c;
      std::ostringstream o;
      o << fr.raw.width << 'x' << fr.raw.height << ':' << int(fr.raw.format) << ':' << int(c.demosaic) << ':'
        << (c.sharpen > 0.0F ? 1 : 0) << ':' << int(c.white_balance) << ':' << int(fr.out.format) << ':'
        << fr.pipeline->params.bins() << ':' << (fr.pipeline->lut.size() + 1 > 4100 ? 1 : 0);
      parts.push_back(o.str());
    }
    std::sort(parts.begin(), parts.end());
    std::ostringstream k;
    k << prop.gcnArchName << '/' << cu_count;
    for (const std::string& p : parts) k << '/' << p;
    return k.str();
  }
};

namespace {

std::string schedule_text(const isp_batch_schedule& s) {
  std::ostringstream o;
  for (const auto& row : s.tile_height) o << row[0] << ' ' << row[1] << ' ';
  for (const auto& row : s.tiles_per_item) o << row[0] << ' ' << row[1] << ' ';
  o << s.awb_tiles_per_item << ' ' << s.blocks_per_cu << ' ' << s.order << ' ' << s.queue << ' ' << s.per_frame;
  return o.str();
}

bool parse_schedule(const std::string& text, isp_batch_schedule* s) {
  std::istringstream in(text);
  isp_batch_schedule r{};
  for (auto& row : r.tile_height) in >> row[0] >> row[1];
  for (auto& row : r.tiles_per_item) in >> row[0] >> row[1];
  in >> r.awb_tiles_per_item >> r.blocks_per_cu >> r.order >> r.queue;
  if (!in) return false;
  if (!(in >> r.per_frame)) r.per_frame = 0;  // entries written before the field existed
  r.graph = s->graph;
  try {
    validate_schedule(r);
  } catch (const std::exception&) {
    return false;
  }
  *s = r;
  return true;
}

// Cache file: one line per entry, "<key>\t<schedule>"; the last entry for a key wins.
bool cache_lookup(const char* path, const std::string& key, isp_batch_schedule* s) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    const std::size_t tab = line.find('\t');
    if (tab == std::string::npos || line.compare(0, tab, key) != 0 || tab != key.size()) continue;
    found = parse_schedule(line.substr(tab + 1), s) || found;
  }
  return found;
}

void cache_store(const char* path, const std::string& key, const isp_batch_schedule& s) {
  std::ofstream out(path, std::ios::app);
  if (!out) throw runtime_failure(std::string("isp_batch_tune: cannot write the cache file ") + path);
  out << key << '\t' << schedule_text(s) << '\n';
}

}  // namespace

extern "C" {

void isp_batch_schedule_init(isp_batch_schedule* schedule) {
  if (schedule != nullptr) schedule_defaults(schedule);
}

isp_status isp_batch_create(isp_batch** out) {
  return guarded([&] {
    if (out == nullptr) throw invalid_argument("out is null");
    *out = nullptr;
    auto b = std::make_unique<isp_batch>();
    schedule_defaults(&b->schedule);
    check(hipGetDevice(&b->device), "hipGetDevice");
    check(hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, b->device),
          "hipDeviceGetAttribute");
    check(hipHostMalloc(reinterpret_cast<void**>(&b->h_table), sizeof(BatchTable), hipHostMallocDefault),
          "hipHostMalloc");
    std::memset(b->h_table, 0, sizeof(BatchTable));
    check(hipMalloc(&b->d_table, sizeof(BatchTable)), "hipMalloc");
    check(hipMalloc(&b->d_control, sizeof(BatchControl)), "hipMalloc");
    check(hipMemset(b->d_control, 0, sizeof(BatchControl)), "hipMemset");
    check(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&b->done, kOrderingEventFlags), "hipEventCreate");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    *out = b.release();
  });
}

isp_status isp_batch_destroy(isp_batch* batch) {
  return guarded([&] { delete batch; });
}

isp_status isp_batch_set_schedule(isp_batch* batch, const isp_batch_schedule* schedule) {
  return guarded([&] {
    if (batch == nullptr || schedule == nullptr) throw invalid_argument("batch or schedule is null");
    validate_schedule(*schedule);
    batch->wait();
    batch->schedule = *schedule;
    if (!batch->frames.empty()) batch->build();
  });
}

isp_status isp_batch_get_schedule(const isp_batch* batch, isp_batch_schedule* schedule) {
  return guarded([&] {
    if (batch == nullptr || schedule == nullptr) throw invalid_argument("batch or schedule is null");
    *schedule = batch->schedule;
  });
}

isp_status isp_batch_set_frames(isp_batch* batch, const isp_batch_frame* frames, int32_t count) {
  return guarded([&] {
    if (batch == nullptr || frames == nullptr) throw invalid_argument("batch or frames is null");
    if (count < 1 || count > ISP_BATCH_MAX_FRAMES) {
      throw invalid_argument("batch frame count must be in 1.." + std::to_string(ISP_BATCH_MAX_FRAMES));
    }
    for (int i = 0; i < count; ++i) {
      const isp_batch_frame& fr = frames[i];
      const std::string at = "batch frame " + std::to_string(i) + ": ";
      if (fr.pipeline == nullptr) throw invalid_argument(at + "pipeline is null");
      for (int j = 0; j < i; ++j) {
        if (frames[j].pipeline == fr.pipeline) throw invalid_argument(at + "its pipeline already has a frame in the batch");
      }
      try {
        validate_raw(fr.raw, fr.pipeline->params.c);
        validate_out(fr.raw, fr.out, fr.pipeline->params.c);
      } catch (const invalid_argument& e) {
        throw invalid_argument(at + e.what());
      }
    }
    batch->wait();  // the previous launch's table copy reads the host table
    batch->frames.assign(frames, frames + count);
    batch->build();
  });
}

isp_status isp_batch_launch(isp_batch* batch, isp_stream_t stream) {
  return guarded([&] {
    if (batch == nullptr) throw invalid_argument("batch is null");
    batch->launch(static_cast<hipStream_t>(stream));
  });
}

isp_status isp_batch_synchronize(isp_batch* batch) {
  return guarded([&] {
    if (batch == nullptr) throw invalid_argument("batch is null");
    batch->wait();
  });
}

isp_status isp_batch_tune(isp_batch* batch, isp_stream_t stream, const char* cache_path, int32_t launches,
                          isp_batch_schedule* chosen, double* best_ms) {
  return guarded([&] {
    if (batch == nullptr) throw invalid_argument("batch is null");
    if (batch->frames.empty()) throw invalid_argument("the batch has no frames (isp_batch_set_frames)");
    if (launches < 1 || launches > 1000) throw invalid_argument("launches must be in 1..1000");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const std::string key = batch->key();
    isp_batch_schedule best = batch->schedule;
    if (cache_path != nullptr && cache_lookup(cache_path, key, &best)) {
      batch->wait();
      batch->schedule = best;
      batch->build();
      if (chosen != nullptr) *chosen = best;
      if (best_ms != nullptr) *best_ms = 0.0;
      return;
    }

    // Save the member streams' state: the timed launches advance it.
    batch->wait();
    for (const isp_batch_frame& fr : batch->frames) {
      if (fr.pipeline->queued) check(hipEventSynchronize(fr.pipeline->done), "hipEventSynchronize");
    }
    const std::size_t n = batch->frames.size();
    std::vector<WbState> wb(n);
    std::vector<AeState> ae(n);
    for (std::size_t i = 0; i < n; ++i) {
      const isp_pipeline* p = batch->frames[i].pipeline;
      check(hipMemcpy(&wb[i], &p->d_scratch->wb, sizeof(WbState), hipMemcpyDeviceToHost), "hipMemcpy");
      if (p->d_ae != nullptr) check(hipMemcpy(&ae[i], p->d_ae, sizeof(AeState), hipMemcpyDeviceToHost), "hipMemcpy");
    }

    hipEvent_t t0 = nullptr;
    hipEvent_t t1 = nullptr;
    check(hipEventCreate(&t0), "hipEventCreate");
    check(hipEventCreate(&t1), "hipEventCreate");
    auto time = [&](const isp_batch_schedule& candidate) {
      batch->wait();
      batch->schedule = candidate;
      batch->build();
      batch->launch(s);  // warm-up (uploads the table, captures the graph)
      check(hipEventRecord(t0, s), "hipEventRecord");
      for (int i = 0; i < launches; ++i) batch->launch(s);
      check(hipEventRecord(t1, s), "hipEventRecord");
      check(hipEventSynchronize(t1), "hipEventSynchronize");
      float ms = 0.0F;
      check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
      return static_cast<double>(ms) / launches;
    };
    double best_time = 0.0;
    try {
      best_time = time(best);
      auto consider = [&](const isp_batch_schedule& candidate) {
        const double t = time(candidate);
        if (t < best_time) {
          best_time = t;
          best = candidate;
        }
      };
      // Per format: the tile height of each (raw format, demosaic) in the batch.
      bool present[4][2] = {};
      bool any_awb = false;
      for (const isp_batch_frame& fr : batch->frames) {
        present[fr.raw.format][fr.pipeline->params.c.demosaic == ISP_DEMOSAIC_MALVAR ? 1 : 0] = true;
        any_awb = any_awb || fr.pipeline->params.c.white_balance != ISP_WB_MANUAL;
      }
      // Per format: the tile height and the work-item size of each (raw format, demosaic).
      for (int fmt = 0; fmt < 4; ++fmt) {
        for (int d = 0; d < 2; ++d) {
          if (!present[fmt][d]) continue;
          isp_batch_schedule c = best;
          c.tile_height[fmt][d] = best.tile_height[fmt][d] == 32 ? 16 : 32;
          consider(c);
          for (int tpi : {1, 2, 4, 8, 16, 32}) {
            if (tpi == best.tiles_per_item[fmt][d]) continue;
            isp_batch_schedule t = best;
            t.tiles_per_item[fmt][d] = tpi;
            consider(t);
          }
        }
      }
      if (any_awb) {
        for (int a : {2, 4, 8, 16, 32, 64}) {
          if (a == best.awb_tiles_per_item) continue;
          isp_batch_schedule c = best;
          c.awb_tiles_per_item = a;
          consider(c);
        }
      }
      {
        isp_batch_schedule c = best;
        c.order = best.order == ISP_BATCH_ORDER_LARGEST_FIRST ? ISP_BATCH_ORDER_AS_GIVEN : ISP_BATCH_ORDER_LARGEST_FIRST;
        consider(c);
      }
      for (int qm : {ISP_BATCH_QUEUE_DYNAMIC, ISP_BATCH_QUEUE_STATIC, ISP_BATCH_QUEUE_STATIC_XCD}) {
        if (qm == best.queue) continue;
        isp_batch_schedule c = best;
        c.queue = qm;
        consider(c);
      }
      {
        const int fit = batch_main_blocks_per_cu(batch->sharpen, batch->wide, batch->uniform);
        for (int bp : {1, 2}) {
          if (bp >= fit || bp == best.blocks_per_cu) continue;
          isp_batch_schedule c = best;
          c.blocks_per_cu = bp;
          consider(c);
        }
      }
      // Last, the per-frame kernels one frame after another: few large frames fill
      // the GPU one at a time, and their kernels hold one frame's tables in the caches.
      {
        isp_batch_schedule c = best;
        c.per_frame = 1;
        consider(c);
      }
    } catch (...) {
      static_cast<void>(hipEventDestroy(t0));
      static_cast<void>(hipEventDestroy(t1));
      throw;
    }
    static_cast<void>(hipEventDestroy(t0));
    static_cast<void>(hipEventDestroy(t1));

    batch->wait();
    batch->schedule = best;
    batch->build();
    for (std::size_t i = 0; i < n; ++i) {
      isp_pipeline* p = batch->frames[i].pipeline;
      upload(&wb[i], sizeof(WbState), &p->d_scratch->wb);
      if (p->d_ae != nullptr) upload(&ae[i], sizeof(AeState), p->d_ae);
    }
    if (cache_path != nullptr) cache_store(cache_path, key, best);
    if (chosen != nullptr) *chosen = best;
    if (best_ms != nullptr) *best_ms = best_time;
  });
}

}  // extern "C"
