// Request server behind vg_server_* (include/vgcut.h): continuous batching of cut requests over
// resident scenes. One worker thread owns the launch side: it completes batches whose event has
// fired, then launches new batches on free lanes, each a group of queued requests of one scene that
// fits the view limit, the scene's free view slots and the HBM budget. Client threads only touch the
// queue and the completed list under the mutex; reads of finished requests use a stream of their own.
//
// Each lane owns its transient buffers (the (view, cell) queue and the work list) and keeps them between
// batches, growing them when a batch needs more: allocating per batch (a stream-ordered pool) cost
// 0.1-0.9 ms of host time per batch, more than a light request's whole cut. The lanes' buffers count
// against the HBM budget; idle lanes give theirs back when a batch cannot otherwise fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vgcut_host.h"
#include "vgcut_kernels.h"

using namespace vgi;

namespace {

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  uint32_t scene = 0;
  std::vector<vg_view> views;
  std::vector<uint32_t> slots;  // assigned view slots, one per view
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = VG_OK;
  uint32_t batch_requests = 0, lane = 0;
  bool done = false;
};

struct SceneEntry {
  vg_scene* scene = nullptr;
  uint64_t resident = 0;
  std::vector<uint32_t> free_slots;
  uint32_t items_per_view = 0;  // most work items per view any batch of this scene used (0: none yet)
  // Estimated run time of one view of this scene: a running mean of (batch run time / views) over the
  // scene's finished batches; before the first one, a guess from the instance count.
  double us_per_view = 0;
  bool measured = false;
};

struct Batch {
  uint32_t scene = 0;
  uint32_t lane = 0;
  uint32_t views = 0;
  std::vector<Request*> requests;
  uint64_t bytes = 0;  // transient bytes the batch needs (its lane holds at least these)
  bool full = false;  // work list at its worst-case size (after an overflow of the optimistic one)
  double t_launch = 0;
};

// A batch's work list starts at this many items (uint4 item + u64 first task + f32 scale: 28 B each)
// when the worst case (every (view, instance, cut)) is larger. Real cuts use a small fraction of the
// worst case; a batch that overflows it is run again at the worst-case size, which gives the same
// results because every request starts from a fresh controller state.
constexpr uint32_t kFirstWorkItems = 1u << 22;
constexpr size_t kWorkItemBytes = sizeof(uint4) + sizeof(unsigned long long) + sizeof(float);

// The first launch of a batch sizes its work list from what batches of the scene used so far (1.25x the
// most items per view, times the views), or `first_items` before any batch ran; never above the worst case.
void batch_sizes(const SceneEntry& se, uint32_t views, bool full, uint32_t first_items, size_t* queue_bytes, size_t* work_items) {
  transient_sizes(se.scene, views, queue_bytes, work_items);
  if (full) return;
  const size_t guess = se.items_per_view ? static_cast<size_t>(se.items_per_view) * views * 5 / 4 + 4096 : first_items;
  // whole partitions (the device splits the list into kWorkParts equal parts)
  *work_items = std::min<size_t>(*work_items, (guess + vg::kWorkParts - 1) / vg::kWorkParts * vg::kWorkParts);
}

// Transient bytes of a launch of `views` views of a scene: the (view, cell) queue and the work list.
uint64_t transient_bytes(const SceneEntry& se, uint32_t views, bool full, uint32_t first_items) {
  size_t queue_bytes = 0, work_items = 0;
  batch_sizes(se, views, full, first_items, &queue_bytes, &work_items);
  return queue_bytes + work_items * kWorkItemBytes;
}

}  // namespace

struct vg_server {
  vg_ctx* ctx = nullptr;
  vg_server_config cfg{};
  std::vector<SceneEntry> scenes;
  std::vector<std::unique_ptr<Lane>> lanes;
  std::vector<hipEvent_t> lane_done;
  std::vector<uint32_t*> lane_flags;  // pinned: the lane's first 8 flag words after its batch
  std::vector<int> lane_batch;        // index into `running` or -1
  hipStream_t read_stream = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  std::deque<uint64_t> completed;
  std::vector<Batch> running;
  std::deque<Batch> deferred;  // reruns at full size waiting for HBM budget (they keep their view slots)
  // `reserved`: bytes of the lanes' transient buffers (lane_bytes per lane), held between batches.
  uint64_t resident = 0, reserved = 0, peak_reserved = 0, batches = 0, requests_done = 0, max_batch_seen = 0, reruns = 0;
  std::vector<uint64_t> lane_bytes;
  uint64_t next_ticket = 1;
  uint64_t generation = 0;  // bumped by submit and release: the worker waits for it when it cannot progress
  bool stop = false;
  std::thread worker;
  std::string error;

  void set_error(const std::string& e) { error = e; }

  // Marks the batches whose event fired as done. Called with the mutex held.
  void complete() {
    for (size_t i = 0; i < running.size();) {
      Batch& b = running[i];
      const hipError_t q = hipEventQuery(lane_done[b.lane]);
      if (q == hipErrorNotReady) {
        i++;
        continue;
      }
      int status = VG_OK;
      if (q != hipSuccess) {
        status = VG_E_DEVICE;
        set_error(std::string("batch failed on the device: ") + hipGetErrorString(q));
      } else if (lane_flags[b.lane][3]) {
        status = VG_E_CAPACITY;
      } else if (lane_flags[b.lane][2]) {
        status = VG_E_DEVICE;
      }
      {
        SceneEntry& se = scenes[b.scene];
        // flags[4]: the most items any partition took; the list is split evenly, so the batch needed
        // kWorkParts times that
        const uint64_t need = static_cast<uint64_t>(lane_flags[b.lane][4]) * vg::kWorkParts;
        const uint32_t per_view = static_cast<uint32_t>((need + b.views - 1) / std::max(1u, b.views));
        se.items_per_view = std::max(se.items_per_view, per_view);
        if (status == VG_OK && b.views) {
          const double run = (now_us() - b.t_launch) / b.views;
          se.us_per_view = se.measured ? 0.75 * se.us_per_view + 0.25 * run : run;
          se.measured = true;
        }
      }
      if (status == VG_E_CAPACITY && !b.full) {
        // The optimistic work list overflowed: run the same batch (same slots) again at full size, on this
        // lane now if its full-size transients fit the budget, else once they do (before new batches).
        Batch again;
        again.scene = b.scene;
        again.views = b.views;
        again.requests = b.requests;
        again.full = true;
        const uint32_t lane = b.lane;
        running.erase(running.begin() + static_cast<std::ptrdiff_t>(i));
        for (size_t k = 0; k < running.size(); k++) lane_batch[running[k].lane] = static_cast<int>(k);
        lane_batch[lane] = -1;
        reruns++;
        if (make_room(lane, transient_bytes(scenes[again.scene], again.views, true, cfg.first_work_items), false))
          launch(again, static_cast<int>(lane));
        else
          deferred.push_back(std::move(again));
        continue;
      }
      const double t = now_us();
      for (Request* r : b.requests) {
        r->t_done = t;
        r->status = status;
        r->done = true;
        completed.push_back(r->ticket);
        requests_done++;
      }
      lane_batch[b.lane] = -1;
      running.erase(running.begin() + static_cast<std::ptrdiff_t>(i));
      for (size_t k = 0; k < running.size(); k++) lane_batch[running[k].lane] = static_cast<int>(k);
      cv_done.notify_all();
    }
  }

  // Frees a lane's transient buffers (the lane is idle: its stream has finished its last batch).
  void shrink(uint32_t lane) {
    Lane* l = lanes[lane].get();
    (void)hipStreamSynchronize(l->stream);
    for (DevBuf* b : {&l->d_cell_queue, &l->d_work, &l->d_work_start, &l->d_work_scale}) {
      if (b->ptr) (void)hipFree(b->ptr);
      *b = DevBuf{};
    }
    l->cell_queue_capacity = 0;
    l->work_capacity = 0;
    reserved -= lane_bytes[lane];
    lane_bytes[lane] = 0;
  }

  // Whether `lane` can hold `need` transient bytes within the HBM budget: it already does, or it may grow
  // into what the other lanes leave; with `shrink_idle`, idle lanes give their buffers back first.
  bool make_room(uint32_t lane, uint64_t need, bool shrink_idle) {
    if (lane_bytes[lane] >= need) return true;
    auto fits = [&] { return resident + (reserved - lane_bytes[lane]) + need <= cfg.hbm_budget; };
    if (fits()) return true;
    if (!shrink_idle) return false;
    for (uint32_t l = 0; l < lanes.size() && !fits(); l++)
      if (l != lane && lane_batch[l] < 0 && lane_bytes[l]) shrink(l);
    return fits();
  }

  // Gives `lane` transient buffers for `queue_bytes` of cell queue and `work_items` work-list items (more
  // when the lane already holds more). The caller made room within the budget. Returns VG_OK or VG_E_NOMEM.
  int grow(uint32_t lane, size_t queue_bytes, size_t work_items) {
    Lane* l = lanes[lane].get();
    const bool queue_ok = l->d_cell_queue.bytes >= queue_bytes, work_ok = l->d_work.bytes >= sizeof(uint4) * work_items;
    if (queue_ok && work_ok) return VG_OK;
    (void)hipStreamSynchronize(l->stream);
    auto realloc = [&](DevBuf& b, size_t bytes) {
      if (b.bytes >= bytes) return true;
      if (b.ptr) (void)hipFree(b.ptr);
      b = DevBuf{};
      void* ptr = nullptr;
      if (hipMalloc(&ptr, std::max<size_t>(bytes, 256)) != hipSuccess) return false;
      b = DevBuf{ptr, std::max<size_t>(bytes, 256)};
      return true;
    };
    const bool ok = realloc(l->d_cell_queue, queue_bytes) && realloc(l->d_work, sizeof(uint4) * work_items) &&
                    realloc(l->d_work_start, sizeof(unsigned long long) * work_items) && realloc(l->d_work_scale, sizeof(float) * work_items);
    const uint64_t now = l->d_cell_queue.bytes + l->d_work.bytes + l->d_work_start.bytes + l->d_work_scale.bytes;
    reserved = reserved - lane_bytes[lane] + now;
    lane_bytes[lane] = now;
    peak_reserved = std::max(peak_reserved, reserved);
    return ok ? VG_OK : VG_E_NOMEM;
  }

  // Launches deferred reruns that fit now, oldest first (idle lanes give their buffers back for them); one
  // that cannot fit even with every other lane idle fails its requests with VG_E_NOMEM. With the mutex held.
  void launch_deferred() {
    while (!deferred.empty()) {
      const int lane = free_lane();
      if (lane < 0) return;
      Batch& b = deferred.front();
      if (!make_room(static_cast<uint32_t>(lane), transient_bytes(scenes[b.scene], b.views, true, cfg.first_work_items), true)) {
        if (!running.empty()) return;  // wait for the lanes in use
        set_error("vg_server: a request's full-size work list does not fit the HBM budget");
        const double t = now_us();
        for (Request* r : b.requests) {
          r->status = VG_E_NOMEM;
          r->done = true;
          r->t_done = t;
          completed.push_back(r->ticket);
          requests_done++;
        }
        deferred.pop_front();
        cv_done.notify_all();
        continue;
      }
      Batch next = std::move(b);
      deferred.pop_front();
      launch(next, lane);
    }
  }

  int free_lane() const {
    for (size_t l = 0; l < lane_batch.size(); l++)
      if (lane_batch[l] < 0) return static_cast<int>(l);
    return -1;
  }

  double estimate(const Request* r) const { return scenes[r->scene].us_per_view * static_cast<double>(r->views.size()); }

  // Picks the next batch (one scene; fits the view limit, the scene's free slots and memory) under the
  // shortest-first policy: the head is the oldest request that has waited longer than max_bypass_us,
  // else the queued request of least estimated run time that fits; the batch is then filled with the
  // head scene's other requests, least estimated run time first. An overdue head that does not fit
  // stops the pick: nothing overtakes it. With the mutex held.
  // What a lane's transient buffers may hold: the budget left by the resident scenes and the other lanes.
  uint64_t lane_room(uint32_t lane) const {
    const uint64_t others = resident + reserved - lane_bytes[lane];
    return cfg.hbm_budget > others ? cfg.hbm_budget - others : 0;
  }

  bool pick_shortest(Batch& b, uint32_t lane) {
    const uint64_t budget_left = lane_room(lane);
    const double t = now_us();
    auto fits = [&](const Request* r, uint32_t views_before) {
      const uint32_t n = views_before + static_cast<uint32_t>(r->views.size());
      return n <= scenes[r->scene].free_slots.size() && n <= cfg.max_batch_views &&
             transient_bytes(scenes[r->scene], n, false, cfg.first_work_items) <= budget_left;
    };
    size_t head = queue.size();
    if (!queue.empty() && t - queue.front()->t_submit > cfg.max_bypass_us) {
      if (!fits(queue.front(), 0)) return false;
      head = 0;
    } else {
      double best = 0;
      for (size_t i = 0; i < queue.size(); i++) {
        const double e = estimate(queue[i]);
        if ((head == queue.size() || e < best) && fits(queue[i], 0)) {
          head = i;
          best = e;
        }
      }
      if (head == queue.size()) return false;
    }
    Request* h = queue[head];
    b.scene = h->scene;
    b.views = static_cast<uint32_t>(h->views.size());
    b.requests.push_back(h);
    std::vector<std::pair<double, Request*>> rest;
    for (Request* r : queue)
      if (r != h && r->scene == b.scene) rest.emplace_back(estimate(r), r);
    std::stable_sort(rest.begin(), rest.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (const auto& e : rest) {
      if (cfg.max_batch_requests && b.requests.size() >= cfg.max_batch_requests) break;
      if (!fits(e.second, b.views)) continue;
      b.views += static_cast<uint32_t>(e.second->views.size());
      b.requests.push_back(e.second);
    }
    return true;
  }

  bool pick(Batch& b, uint32_t lane) { return cfg.policy == VG_SERVE_FIFO ? pick_fifo(b, lane) : pick_shortest(b, lane); }

  // Picks the next batch from the queue (FIFO, one scene, fits views, slots and memory). With the mutex held.
  bool pick_fifo(Batch& b, uint32_t lane) {
    const uint64_t budget_left = lane_room(lane);
    const double t = now_us();
    std::vector<uint32_t> slots_left(scenes.size());
    for (size_t k = 0; k < scenes.size(); k++) slots_left[k] = static_cast<uint32_t>(scenes[k].free_slots.size());
    bool have = false;
    for (size_t i = 0; i < queue.size(); i++) {
      Request* r = queue[i];
      const SceneEntry& se = scenes[r->scene];
      const uint32_t n = static_cast<uint32_t>(r->views.size());
      if (!have) {
        if (n <= slots_left[r->scene] && n <= cfg.max_batch_views && transient_bytes(se, n, false, cfg.first_work_items) <= budget_left) {
          b.scene = r->scene;
          b.views = n;
          b.requests.push_back(r);
          have = true;
        } else if (i == 0 && t - r->t_submit > cfg.max_bypass_us) {
          return false;  // the oldest request has waited long enough: nothing overtakes it
        }
        continue;
      }
      if (r->scene != b.scene) continue;
      if (cfg.max_batch_requests && b.requests.size() >= cfg.max_batch_requests) break;
      if (b.views + n > cfg.max_batch_views || b.views + n > slots_left[b.scene]) continue;
      if (transient_bytes(scenes[b.scene], b.views + n, false, cfg.first_work_items) > budget_left) continue;
      b.views += n;
      b.requests.push_back(r);
    }
    return have;
  }

  // Launches a picked batch on a lane. With the mutex held.
  int launch(Batch& b, int lane_index) {
    SceneEntry& se = scenes[b.scene];
    vg_scene* s = se.scene;
    Lane* lane = lanes[lane_index].get();
    std::vector<vg_view> views;
    views.reserve(b.views);
    const double t = now_us();
    for (Request* r : b.requests) {
      if (!b.full) {  // a first launch: assign view slots and leave the queue (a rerun keeps both)
        r->slots.clear();
        for (size_t i = 0; i < r->views.size(); i++) {
          r->slots.push_back(se.free_slots.back());
          se.free_slots.pop_back();
        }
        r->t_start = t;
        queue.erase(std::find(queue.begin(), queue.end(), r));
      }
      for (size_t i = 0; i < r->views.size(); i++) {
        views.push_back(r->views[i]);
        views.back().slot = r->slots[i];
      }
      r->batch_requests = static_cast<uint32_t>(b.requests.size());
      r->lane = static_cast<uint32_t>(lane_index);
    }
    size_t queue_bytes = 0, work_items = 0;
    batch_sizes(se, b.views, b.full, cfg.first_work_items, &queue_bytes, &work_items);
    b.bytes = transient_bytes(se, b.views, b.full, cfg.first_work_items);
    b.lane = static_cast<uint32_t>(lane_index);
    b.t_launch = t;
    int rc = grow(static_cast<uint32_t>(lane_index), queue_bytes, work_items);
    if (rc == VG_OK) {
      // The whole work list the lane holds (whole partitions): a batch runs with at least what it asked for.
      lane->borrowed = true;
      lane->cell_queue_capacity = lane->d_cell_queue.bytes / sizeof(uint2);
      const size_t items = lane->d_work.bytes / sizeof(uint4);
      size_t full_queue = 0, full_items = 0;
      transient_sizes(s, b.views, &full_queue, &full_items);
      lane->work_capacity = static_cast<uint32_t>(std::min(items, full_items) / vg::kWorkParts * vg::kWorkParts);
      rc = prepare_cut(s, lane, views.data(), b.views, true);
      if (rc == VG_OK) rc = enqueue_cut(s, lane, b.views, false);
      if (rc == VG_OK && hipMemcpyAsync(lane_flags[lane_index], lane->d_flags.ptr, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                        lane->stream) != hipSuccess)
        rc = VG_E_DEVICE;
    }
    if (rc == VG_OK && hipEventRecord(lane_done[lane_index], lane->stream) != hipSuccess) rc = VG_E_DEVICE;
    if (rc != VG_OK) {
      set_error("batch launch failed: " + ctx->error);
      (void)hipStreamSynchronize(lane->stream);
      const double t_fail = now_us();
      for (Request* r : b.requests) {
        r->status = rc;
        r->done = true;
        r->t_done = t_fail;
        completed.push_back(r->ticket);
        requests_done++;
      }
      cv_done.notify_all();
      return rc;
    }
    batches++;
    max_batch_seen = std::max<uint64_t>(max_batch_seen, b.requests.size());
    running.push_back(std::move(b));
    lane_batch[lane_index] = static_cast<int>(running.size() - 1);
    return VG_OK;
  }

  void run() {
    if (hipSetDevice(ctx->cfg.device) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu);
      set_error("hipSetDevice failed in the server worker");
    }
    std::unique_lock<std::mutex> lk(mu);
    uint64_t seen = generation;
    while (!stop) {
      complete();
      const size_t waiting = deferred.size();
      launch_deferred();
      bool launched = deferred.size() != waiting;
      for (int lane = free_lane(); lane >= 0 && deferred.empty() && !queue.empty(); lane = free_lane()) {
        Batch b;
        if (!pick(b, static_cast<uint32_t>(lane))) break;
        launch(b, lane);
        launched = true;
      }
      if (running.empty() && !launched) {
        // Nothing on the device and nothing launchable (empty queue, or every queued request waits for
        // view slots its client has not released): sleep until a submit or a release changes that.
        cv_work.wait(lk, [&] { return stop || generation != seen; });
        seen = generation;
        continue;
      }
      // Batches in flight: poll their events, letting clients in between.
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    }
  }
};

extern "C" {

int vg_server_create(vg_ctx* ctx, const vg_server_config* config, vg_server** out) {
  if (!ctx || !out) return VG_E_INVALID;
  *out = nullptr;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "vg_server_create: host-only context (device < 0)");
  vg_server_config c{};
  if (config) c = *config;
  if (!c.lanes) c.lanes = 4;
  if (!c.max_batch_views) c.max_batch_views = ctx->cfg.max_views;
  if (!c.max_bypass_us) c.max_bypass_us = c.policy == VG_SERVE_FIFO ? 2000 : 20000;
  if (c.policy > VG_SERVE_FIFO) return fail(ctx, VG_E_INVALID, "vg_server_create: unknown policy");
  if (!c.first_work_items) c.first_work_items = kFirstWorkItems;
  if (c.lanes > 16 || c.max_batch_views > ctx->cfg.max_views) return fail(ctx, VG_E_INVALID, "vg_server_create: lanes 1-16, max_batch_views <= max_views");
  if (!c.hbm_budget) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(ctx, VG_E_DEVICE, "vg_server_create: hipMemGetInfo");
    c.hbm_budget = static_cast<uint64_t>(0.9 * static_cast<double>(total_b));
  }
  auto* sv = new (std::nothrow) vg_server();
  if (!sv) return fail(ctx, VG_E_NOMEM, "out of host memory");
  sv->ctx = ctx;
  sv->cfg = c;
  int rc = VG_OK;
  auto init = [&]() -> int {
    if (hipStreamCreateWithFlags(&sv->read_stream, hipStreamNonBlocking) != hipSuccess) return fail(ctx, VG_E_DEVICE, "hipStreamCreate");
    for (uint32_t l = 0; l < c.lanes; l++) {
      sv->lanes.emplace_back(new Lane());
      if (int r = lane_init(ctx, sv->lanes.back().get(), nullptr)) return r;
      if (int r = lane_reserve_views(ctx, sv->lanes.back().get())) return r;
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(ctx, VG_E_DEVICE, "hipEventCreate");
      sv->lane_done.push_back(e);
      uint32_t* flags = nullptr;
      if (hipHostMalloc(reinterpret_cast<void**>(&flags), 8 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return fail(ctx, VG_E_NOMEM, "hipHostMalloc");
      std::memset(flags, 0, 8 * sizeof(uint32_t));
      sv->lane_flags.push_back(flags);
      sv->lane_batch.push_back(-1);
      sv->lane_bytes.push_back(0);
      sv->lanes.back()->max_blocks_per_cu = c.max_blocks_per_cu;
    }
    return VG_OK;
  };
  if ((rc = init())) {
    vg_server_destroy(sv);
    return rc;
  }
  sv->worker = std::thread([sv] { sv->run(); });
  *out = sv;
  return VG_OK;
}

void vg_server_destroy(vg_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->generation++;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto& lane : sv->lanes) {
    if (lane->stream) (void)hipStreamSynchronize(lane->stream);
    for (DevBuf* b : {&lane->d_cell_queue, &lane->d_work, &lane->d_work_start, &lane->d_work_scale})
      if (b->ptr) (void)hipFree(b->ptr);
    lane->d_cell_queue = lane->d_work = lane->d_work_start = lane->d_work_scale = DevBuf{};
    lane->borrowed = true;  // the server freed its transient buffers above
    lane_destroy(lane.get());
  }
  for (hipEvent_t e : sv->lane_done) (void)hipEventDestroy(e);
  for (uint32_t* f : sv->lane_flags) (void)hipHostFree(f);
  if (sv->read_stream) (void)hipStreamDestroy(sv->read_stream);
  delete sv;
}

const char* vg_server_last_error(const vg_server* sv) { return sv ? sv->error.c_str() : "no server"; }

int vg_server_add_scene(vg_server* sv, vg_scene* s, uint32_t* scene_id) {
  if (!sv || !s || !scene_id || s->ctx != sv->ctx) return VG_E_INVALID;
  vg_ctx* ctx = sv->ctx;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (int rc = upload_scene(s)) return rc;
  for (uint32_t slot = 0; slot < ctx->cfg.max_views; slot++)
    if (int rc = ensure_slot(s, slot)) return rc;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, VG_E_DEVICE, "vg_server_add_scene: upload");
  const uint64_t bytes = vg_scene_bytes(s);
  if (sv->resident + bytes > sv->cfg.hbm_budget) {
    sv->set_error("vg_server_add_scene: the scene does not fit the HBM budget");
    return VG_E_NOMEM;
  }
  SceneEntry e;
  e.scene = s;
  e.resident = bytes;
  e.us_per_view = 20.0 + 1e-4 * static_cast<double>(s->i_slot.size());  // replaced by the first measured batch
  for (uint32_t slot = ctx->cfg.max_views; slot-- > 0;) e.free_slots.push_back(slot);
  sv->resident += bytes;
  sv->scenes.push_back(std::move(e));
  *scene_id = static_cast<uint32_t>(sv->scenes.size() - 1);
  return VG_OK;
}

int vg_server_submit(vg_server* sv, uint32_t scene_id, const vg_view* views, uint32_t count, uint64_t* ticket) {
  if (!sv || !views || !count || !ticket) return VG_E_INVALID;
  std::unique_ptr<Request> r(new (std::nothrow) Request());
  if (!r) return VG_E_NOMEM;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (scene_id >= sv->scenes.size()) return VG_E_INVALID;
    const SceneEntry& se = sv->scenes[scene_id];
    if (count > sv->cfg.max_batch_views || count > sv->ctx->cfg.max_views) {
      sv->set_error("vg_server_submit: more views than a batch holds");
      return VG_E_INVALID;
    }
    if (sv->resident + transient_bytes(se, count, false, sv->cfg.first_work_items) > sv->cfg.hbm_budget) {
      sv->set_error("vg_server_submit: the request alone does not fit the HBM budget");
      return VG_E_NOMEM;
    }
    for (uint32_t i = 0; i < count; i++)
      if (!(views[i].proj_scale > 0) || !(views[i].near_plane > 0) ||
          (views[i].triangle_budget && !(views[i].threshold_range[0] > 0 && views[i].threshold_range[1] >= views[i].threshold_range[0]))) {
        sv->set_error("vg_server_submit: invalid view");
        return VG_E_INVALID;
      }
    r->ticket = sv->next_ticket++;
    r->scene = scene_id;
    r->views.assign(views, views + count);
    r->t_submit = now_us();
    *ticket = r->ticket;
    sv->queue.push_back(r.get());
    sv->requests.emplace(r->ticket, std::move(r));
    sv->generation++;
  }
  sv->cv_work.notify_one();
  return VG_OK;
}

int vg_server_poll(vg_server* sv, vg_request_done* done, uint32_t max_done, uint32_t* count, uint32_t wait_us) {
  if (!sv || !count || (max_done && !done)) return VG_E_INVALID;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->completed.empty() && wait_us)
    sv->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !sv->completed.empty(); });
  uint32_t n = 0;
  while (n < max_done && !sv->completed.empty()) {
    const uint64_t t = sv->completed.front();
    sv->completed.pop_front();
    auto it = sv->requests.find(t);
    if (it == sv->requests.end()) continue;
    const Request& r = *it->second;
    vg_request_done& d = done[n++];
    d.ticket = r.ticket;
    d.status = r.status;
    d.views = static_cast<uint32_t>(r.views.size());
    d.batch_requests = r.batch_requests;
    d.lane = r.lane;
    d.queued_us = r.t_start - r.t_submit;
    d.run_us = r.t_done - r.t_start;
    d.total_us = r.t_done - r.t_submit;
  }
  *count = n;
  return VG_OK;
}

int vg_server_read(vg_server* sv, uint64_t ticket, uint32_t view_index, uint32_t* counters, uint32_t* draw_args, uint32_t* draw_list,
                   uint32_t max_entries, vg_metrics* metrics) {
  if (!sv) return VG_E_INVALID;
  vg_scene* s = nullptr;
  uint32_t slot = 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->requests.find(ticket);
    if (it == sv->requests.end() || !it->second->done || it->second->status != VG_OK || view_index >= it->second->slots.size())
      return VG_E_INVALID;
    s = sv->scenes[it->second->scene].scene;
    slot = it->second->slots[view_index];
  }
  return read_slot(s, sv->read_stream, slot, counters, draw_args, draw_list, max_entries, metrics);
}

int vg_server_release(vg_server* sv, uint64_t ticket) {
  if (!sv) return VG_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->requests.find(ticket);
  if (it == sv->requests.end() || !it->second->done) return VG_E_INVALID;
  SceneEntry& se = sv->scenes[it->second->scene];
  for (uint32_t slot : it->second->slots) se.free_slots.push_back(slot);
  sv->requests.erase(it);
  sv->generation++;
  sv->cv_work.notify_one();  // freed slots may let a queued request in
  return VG_OK;
}

int vg_server_stats_get(const vg_server* sv, vg_server_stats* st) {
  if (!sv || !st) return VG_E_INVALID;
  auto* m = const_cast<vg_server*>(sv);
  std::lock_guard<std::mutex> lk(m->mu);
  st->hbm_budget = sv->cfg.hbm_budget;
  st->resident_bytes = sv->resident;
  st->inflight_bytes = sv->reserved;
  st->peak_inflight_bytes = sv->peak_reserved;
  st->batches = sv->batches;
  st->requests_done = sv->requests_done;
  st->max_batch_requests_seen = sv->max_batch_seen;
  st->reruns = sv->reruns;
  st->queued = static_cast<uint32_t>(sv->queue.size());
  st->running = 0;
  for (const Batch& b : sv->running) st->running += static_cast<uint32_t>(b.requests.size());
  return VG_OK;
}

}  // extern "C"
