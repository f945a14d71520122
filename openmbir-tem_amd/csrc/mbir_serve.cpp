// mbir_serve.cpp -- request server: reconstruction jobs arrive at any time, are
// admitted against an HBM budget, and run their execute() loops concurrently on
// one GPU with continuous batching at ICD-step granularity.
//
// Each "tick" is one launch per kernel shape that carries the next conflict-free
// step of every resident job, so jobs of any size share launches, a job that
// arrives joins at the next tick, and a job that converges leaves at once.  Every
// job keeps its volume, error sinogram, column table and nuisance parameters on
// the device between sweeps.  A job's steps, sweeps and nuisance updates are the
// same operations in the same order as mbir_job_run, so its result does not
// depend on what else is running.
//
// Two launch lanes: jobs whose sweeps have many steps (at least long_steps
// conflict-free steps, N_z S) run their ticks on a stream of their own, so that a
// long job's step is not held behind a full launch of the others: the two lanes'
// launches run side by side on the device, and the long job, whose chain of steps
// bounds its latency, keeps close to its own pace.  The lane changes only where a
// job's launches go, not what they compute.
//
// Threads: setup workers admit queued jobs in arrival order (FIFO, no overtaking)
// when the job fits the budget and a slot is free, and run the job's setup
// (mbir_job_create) on the job's own stream; one tick thread owns the launch
// stream, reads each sweep's accumulators back asynchronously (pinned memory +
// event), applies the stop test and drives the outer/inner loop of
// HAADF/HAADF_ReconstructionEngine.cpp:706-815.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "mbir_host.h"

using namespace mbir;

namespace {

constexpr int kMaxServerJobs = 256;
constexpr int kTickEvery = 4;   // ticks per run-ahead event
constexpr int kTickEvents = 4;  // events in flight: the host runs at most 16 ticks ahead
constexpr int kLanes = 2;       // 0: every job; 1: jobs with long sweeps
constexpr int kLongStepsDefault = 2048;
constexpr double kDefBragg = 5e10;  // DefBraggThreshold (BrightField/BFConstants.h)
constexpr int kLongJobsDefault = 4;

struct Req {
  uint64_t ticket = 0;
  mbir_problem prob{};  // host pointers into the copies below until setup is done
  std::vector<double> angles, gain, offset, variance;
  std::vector<float> counts, init;
  int state = MBIR_JOB_QUEUED;
  int error = 0;
  mbir_job* job = nullptr;
  int slot = -1;
  int64_t est = 0, reserved = 0;
  // position in execute()'s loop
  int outer = 0, inner = 0, num_iter = 0, status = 1, sweeps = 0, outer_done = 0;
  int step = 0, nsteps = 0;
  bool waiting = false;    // the sweep's last step is launched; accumulators in flight
  bool finishing = false;  // final nuisance update in flight
  int ypl = 1, nv = 1;
  int lane = 0;
  bool fw = false;          // bright field (per-entry weights)
  double bragg_tmp = 0.0;   // bright field: the threshold of the sweeps to come
  SweepAccum* hacc = nullptr;  // pinned
  LineRef* hsched = nullptr;   // pinned copy of the current sweep's schedule
  hipEvent_t ev = nullptr;
  int64_t visits = 0, updated = 0;
  double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
  std::atomic<bool> cancel{false};
};

int64_t estimate_bytes(const mbir_problem& p)
{
  const mbir_geometry& g = p.geom;
  const int64_t V = (int64_t)g.N_z * g.N_x * g.N_y;
  const int64_t M = (int64_t)g.N_theta * g.N_r * g.N_t;
  const int64_t L = (int64_t)g.N_z * g.N_x;
  const int64_t NB = (int64_t)std::ceil(2.0 * g.delta_xz / g.delta_r - 1e-12) + 1;
  // x; e, y; FP64 projection during setup; per-line arrays; column table bound;
  // bright field: the per-entry weights and the counts kept for the noise step
  return V * 4 + M * 8 + M * 8 + L * 21 + L * g.N_theta * NB * 16 + (1 << 20) +
         (p.bf_mode ? M * 8 : 0);
}

// The job's descriptor for the server's launches: the plain error-sinogram update on
// conflict-free schedules, fp32 atomics on the relaxed one (as the job's own path).
DevJob tick_desc(const mbir_job* j)
{
  DevJob d = j->J;
  d.e_atomic = j->prob.schedule == 1 ? 1 : 0;
  return d;
}

}  // namespace

struct mbir_server {
  mbir_ctx* ctx = nullptr;
  mbir::LineTuning tune;  // the context's kernel-form choice when the server was created
  int max_jobs = 64;
  int long_steps = kLongStepsDefault, long_jobs = kLongJobsDefault;
  uint64_t budget = 0, reserved = 0, peak_reserved = 0;
  int resident = 0;  // jobs holding a slot (in setup, running or finished but unreleased)
  hipStream_t lane_stream[kLanes] = {nullptr, nullptr};
  DevBuf<DevJob> djobs;
  std::vector<int> free_slots;
  std::mutex mu;
  std::condition_variable cv_setup, cv_run, cv_done;
  std::deque<Req*> queue;
  std::vector<Req*> ready;
  std::map<uint64_t, Req*> reqs;
  std::deque<uint64_t> done_fifo;
  uint64_t next_ticket = 1;
  bool stopping = false;
  int setups_running = 0;
  int fatal = 0;
  std::vector<std::thread> setup_thr;
  std::thread run_thr;
  std::chrono::steady_clock::time_point t0;
  mbir_server_stats st{};

  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  hipStream_t stream_of(const Req* r) const { return lane_stream[r->lane]; }
  void setup_loop();
  void run_loop();
  int start_sweep(Req* r);
  int end_sweep(Req* r);
  void finish(Req* r, int state, int error);  // with mu held
};

// ---------------------------------------------------------------------------
void mbir_server::setup_loop()
{
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_setup.wait(lk, [&] {
      if (stopping) return true;
      if (queue.empty()) return false;
      Req* h = queue.front();
      if ((uint64_t)h->est > budget) return true;  // can never fit: fail it
      return resident < max_jobs && reserved + (uint64_t)h->est <= budget;
    });
    if (stopping) return;
    Req* r = queue.front();
    queue.pop_front();
    if ((uint64_t)r->est > budget) {
      finish(r, MBIR_JOB_FAILED, MBIR_ERR_BUDGET);
      continue;
    }
    r->reserved = r->est;
    reserved += r->est;
    peak_reserved = std::max(peak_reserved, reserved);
    resident++;
    setups_running++;
    r->state = MBIR_JOB_SETUP;
    r->t_admit = now();
    lk.unlock();

    mbir_job* job = nullptr;
    int rc = mbir_job_create(ctx, &r->prob, &job);
    const size_t L = rc == 0 ? job->sched_cap() : 0;
    if (rc == 0 && (hipHostMalloc((void**)&r->hacc, sizeof(SweepAccum)) != hipSuccess ||
                    hipHostMalloc((void**)&r->hsched, std::max<size_t>(L, 1) * sizeof(LineRef)) !=
                        hipSuccess ||
                    hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess))
      rc = MBIR_ERR_HIP;
    // bright field: one detector row per slice and full sweeps (NHICD lists run
    // as single jobs)
    if (rc == 0 && (job->J.unit_max > 128 || job->J.vlr_max > 2 ||
                    (job->J.wfull && (job->prob.bf_nhicd || job->J.vlr_max > 1))))
      rc = MBIR_ERR_UNSUPPORTED;

    lk.lock();
    setups_running--;
    if (rc != 0) {
      if (job) mbir_job_destroy(job);
      reserved -= (uint64_t)r->reserved;
      r->reserved = 0;
      resident--;
      finish(r, MBIR_JOB_FAILED, rc);
      cv_setup.notify_all();
      continue;
    }
    r->job = job;
    // the measured footprint replaces the estimate
    reserved = reserved - (uint64_t)r->reserved + (uint64_t)job->bytes;
    r->reserved = job->bytes;
    peak_reserved = std::max(peak_reserved, reserved);
    std::vector<double>().swap(r->angles);
    std::vector<float>().swap(r->counts);
    std::vector<float>().swap(r->init);
    ready.push_back(r);
    cv_run.notify_all();
    cv_setup.notify_all();
  }
}

// Leaves the tick loop: records the outcome and wakes waiters (mu held).
void mbir_server::finish(Req* r, int state, int error)
{
  r->state = state;
  r->error = error;
  r->t_done = now();
  if (r->slot >= 0) {
    free_slots.push_back(r->slot);
    r->slot = -1;
  }
  if (state == MBIR_JOB_DONE) st.jobs_done++;
  else st.jobs_failed++;
  done_fifo.push_back(r->ticket);
  cv_done.notify_all();
}

// Schedule of the job's next sweep: seeded step order, uploaded from pinned
// memory on the launch stream.
int mbir_server::start_sweep(Req* r)
{
  mbir_job* j = r->job;
  hipStream_t stream = stream_of(r);
  if (r->fw) {
    // the threshold 5e10 for the inner sweeps of an outer iteration with more than
    // one (BrightField/BFReconstructionEngine.cpp:655-660); the descriptor carries it
    if (r->num_iter > 1 && r->inner == 0) set_bragg(j, kDefBragg);
    HIP_TRY(launch_set_devjob(djobs.p, r->slot, tick_desc(j), stream));
  }
  build_schedule(j, true);
  if (!j->sched_h.empty())
    std::memcpy(r->hsched, j->sched_h.data(), j->sched_h.size() * sizeof(LineRef));
  int rc = prepare_sweep_on(j, stream, r->hsched);
  if (rc < 0) return rc;
  r->step = 0;
  r->nsteps = (int)j->step_ptr.size() - 1;
  r->waiting = false;
  if (r->nsteps == 0) {  // nothing to update: the sweep ends at once
    HIP_TRY(launch_sweep_reduce(j->sched.p, j->ustat.p, 0, j->acc.p, j->red.p, stream));
    HIP_TRY(hipMemcpyAsync(r->hacc, j->acc.p, sizeof(SweepAccum), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(r->ev, stream));
    r->waiting = true;
  }
  return MBIR_OK;
}

// A sweep's accumulators are back: stop test, then the loop of
// HAADF/HAADF_ReconstructionEngine.cpp:706-815 (mbir_job_run is the same loop).
// Returns 1 when the job has left the tick loop.
int mbir_server::end_sweep(Req* r)
{
  mbir_job* j = r->job;
  hipStream_t stream = stream_of(r);
  mbir_sweep_stats ss{};
  const int status = sweep_status(j, *r->hacc, 0.0, &ss);
  r->waiting = false;
  r->sweeps++;
  r->status = status;
  r->visits += ss.updated + ss.zero_skipped;
  r->updated += ss.updated;
  if (r->cancel.load() || j->cancel.load()) {
    std::lock_guard<std::mutex> lk(mu);
    finish(r, MBIR_JOB_FAILED, MBIR_ERR_CANCELLED);
    return 1;
  }
  if (r->fw) {
    // bright field (BrightField/BFReconstructionEngine.cpp:641-870; mbir_job_run is
    // the same loop): T from the last inner sweep of a multi-sweep loop on; the stop
    // test ends the run before the nuisance step, which runs only for more than one
    // outer iteration
    const double des = j->prob.bragg_T;
    if (status != 0 && r->inner == r->num_iter - 1 && r->inner != 0) {
      r->bragg_tmp = des;
      set_bragg(j, r->bragg_tmp);
    }
    r->inner++;
    if (status != 0 && r->inner < r->num_iter) return start_sweep(r) < 0 ? -1 : 0;
    const int outer = r->outer;
    r->outer++;
    r->outer_done = r->outer;
    const bool stop = status == 0 && outer >= 1;
    if (!stop && j->prob.num_outer > 1 && (j->prob.joint_estimation || j->prob.noise_model)) {
      if (nuisance_update_on(j, stream) < 0) return -1;
      if (launch_set_devjob(djobs.p, r->slot, tick_desc(j), stream) != hipSuccess) return -1;
    }
    if (stop || r->outer >= j->prob.num_outer) {
      if (hipEventRecord(r->ev, stream) != hipSuccess) return -1;
      r->finishing = true;
      return 0;
    }
    r->num_iter = 1;
    r->inner = 0;
    set_bragg(j, r->bragg_tmp);
    return start_sweep(r) < 0 ? -1 : 0;
  }
  r->inner++;
  if (status != 0 && r->inner < r->num_iter) return start_sweep(r) < 0 ? -1 : 0;
  // end of this outer iteration's inner loop
  const int outer = r->outer;
  if (j->prob.joint_estimation || j->prob.noise_model) {
    if (nuisance_update_on(j, stream) < 0) return -1;
    if (launch_set_devjob(djobs.p, r->slot, tick_desc(j), stream) != hipSuccess) return -1;
  }
  r->outer++;
  r->outer_done = r->outer;
  if ((status == 0 && outer >= 1) || r->outer >= j->prob.num_outer) {
    // done once the stream has passed the last nuisance update
    if (hipEventRecord(r->ev, stream) != hipSuccess) return -1;
    r->finishing = true;
    return 0;
  }
  r->num_iter = 1;  // :712-715
  r->inner = 0;
  return start_sweep(r) < 0 ? -1 : 0;
}

void mbir_server::run_loop()
{
  (void)hipSetDevice(ctx->device);
  std::vector<Req*> act;
  // per lane: run-ahead events (one every kTickEvery ticks) and the tick count
  hipEvent_t tick_ev[kLanes][kTickEvents];
  int64_t lane_ticks[kLanes] = {0, 0};
  for (auto& l : tick_ev)
    for (auto& e : l) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
  int64_t tick = 0, n_launch = 0, n_entries = 0, n_wg = 0;
  int long_resident = 0;
  TickArgs args;
  auto fail_all = [&](int code) {
    std::lock_guard<std::mutex> lk(mu);
    fatal = code;
    for (Req* r : act) finish(r, MBIR_JOB_FAILED, code);
    act.clear();
    long_resident = 0;
  };
  // whether lane l may launch another tick: the host stays at most
  // kTickEvery * kTickEvents ticks ahead of the device on each lane
  auto lane_open = [&](int l) {
    const int64_t t = lane_ticks[l];
    if (t % kTickEvery != 0 || t / kTickEvery < kTickEvents) return true;
    return hipEventQuery(tick_ev[l][(t / kTickEvery) % kTickEvents]) != hipErrorNotReady;
  };
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      if (act.empty())
        cv_run.wait(lk, [&] {
          return !ready.empty() || (stopping && queue.empty() && setups_running == 0);
        });
      if (act.empty() && ready.empty() && stopping && queue.empty() && setups_running == 0) break;
      for (Req* r : ready) {
        r->slot = free_slots.back();
        free_slots.pop_back();
        r->state = MBIR_JOB_RUNNING;
        r->t_start = now();
        r->num_iter = r->job->prob.num_inner;
        icd_shape(r->job->J, r->job->J.vlr_max, &r->ypl, &r->nv);
        r->fw = r->job->J.wfull != nullptr;
        if (r->fw) {
          r->nv = 1;  // one detector row per slice
          r->bragg_tmp = r->num_iter > 1 ? kDefBragg : r->job->prob.bragg_T;
          set_bragg(r->job, r->bragg_tmp);
        }
        // the lane: jobs with long sweeps (conflict-free steps per sweep) go to lane 1
        const int64_t spp = (int64_t)r->job->g.N_z * r->job->spacing;
        r->lane = (long_steps > 0 && spp >= long_steps && long_resident < long_jobs) ? 1 : 0;
        long_resident += r->lane;
        st.long_lane_jobs += r->lane;
        act.push_back(r);
      }
      st.peak_resident = std::max<int32_t>(st.peak_resident, (int32_t)act.size());
      std::vector<Req*> fresh;
      fresh.swap(ready);
      lk.unlock();
      bool bad = false;
      for (Req* r : fresh) {
        if (launch_set_devjob(djobs.p, r->slot, tick_desc(r->job), stream_of(r)) != hipSuccess ||
            start_sweep(r) < 0)
          bad = true;
      }
      if (bad) {
        fail_all(MBIR_ERR_HIP);
        continue;
      }
    }
    // sweeps whose accumulators (or final nuisance updates) have arrived
    bool bad = false;
    for (Req* r : act) {
      if (!r->waiting && !r->finishing) continue;
      const hipError_t q = hipEventQuery(r->ev);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) {
        bad = true;
        break;
      }
      if (r->finishing) {
        r->finishing = false;
        long_resident -= r->lane;
        std::lock_guard<std::mutex> lk(mu);
        finish(r, MBIR_JOB_DONE, 0);
        continue;
      }
      if (end_sweep(r) < 0) {
        bad = true;
        break;
      }
      if (r->state == MBIR_JOB_FAILED) long_resident -= r->lane;
    }
    if (bad) {
      fail_all(MBIR_ERR_HIP);
      continue;
    }
    act.erase(std::remove_if(act.begin(), act.end(),
                             [](Req* r) {
                               return r->state == MBIR_JOB_DONE || r->state == MBIR_JOB_FAILED;
                             }),
              act.end());
    // one tick per lane that may run ahead: the next step of every job of the lane
    // that has one, one launch per kernel shape
    bool launched_any = false, held = false;
    bool lane_has[kLanes] = {false, false};
    for (Req* r : act)
      if (!r->waiting && !r->finishing && r->step < r->nsteps) lane_has[r->lane] = true;
    for (int lane = 0; lane < kLanes && !bad; ++lane) {
      if (!lane_has[lane]) continue;
      // a job alone in a launch takes the plain path's step kernel only while the other
      // lane is idle: beside a full launch of the other lane, the batched form (fewer
      // registers) shares the compute units better
      const bool alone_ok = !lane_has[1 - lane];
      if (!lane_open(lane)) {
        held = true;
        continue;
      }
      hipStream_t stream = lane_stream[lane];
      bool launched = false;
      for (int shape = 0; shape < 8 && !bad; ++shape) {
        const int ypl = 1 + ((shape >> 1) & 1), nv = 1 + (shape & 1), fw = shape >> 2;
        if (fw && nv != 1) continue;
        args.n = 0;
        args.slots = 4;
        int wg = 0;
        std::vector<Req*> last;  // jobs whose sweep ends with this launch
        Req* first = nullptr;    // the launch's first job
        auto flush = [&]() -> bool {
          if (args.n == 0) return true;
          if (args.n == 1 && alone_ok) {
            // one job alone in the launch with the other lane idle: the plain path's
            // step kernel, with the job descriptor and the step's units in the launch
            // arguments (no entry search, no descriptor or unit loads before the
            // line's own)
            const TickEntry& E = args.e[0];
            mbir_job* j = first->job;
            if (launch_icd_step(tick_desc(j), j->sched.p + E.line0, E.nlines, E.line0, (E.flags & 1) ? 0 : 1,
                                0, (E.flags >> 1) & 1, stream, j->sched_h.data() + E.line0) !=
                hipSuccess)
              return false;
          } else if (launch_icd_tick(ypl, nv, djobs.p, args, wg, tune.choose(ypl, nv, wg), stream,
                                     fw) != hipSuccess) {
            return false;
          }
          n_launch++;
          n_entries += args.n;
          n_wg += wg;
          launched = true;
          for (Req* r : last) {
            mbir_job* lj = r->job;
            if (launch_sweep_reduce(lj->sched.p, lj->ustat.p, (int)lj->sched_h.size(), lj->acc.p,
                                    lj->red.p, stream) != hipSuccess ||
                hipMemcpyAsync(r->hacc, lj->acc.p, sizeof(SweepAccum), hipMemcpyDeviceToHost,
                               stream) != hipSuccess ||
                hipEventRecord(r->ev, stream) != hipSuccess)
              return false;
            r->waiting = true;
          }
          last.clear();
          args.n = 0;
          args.slots = 4;
          wg = 0;
          return true;
        };
        for (Req* r : act) {
          if (r->lane != lane || r->ypl != ypl || r->nv != nv || (int)r->fw != fw || r->waiting ||
              r->finishing || r->step >= r->nsteps)
            continue;
          mbir_job* j = r->job;
          if (args.n == 0) first = r;
          TickEntry& E = args.e[args.n++];
          args.slots = std::max(args.slots, j->J.col_slots);
          E.job = r->slot;
          E.line0 = j->step_ptr[r->step];
          E.nlines = j->step_ptr[r->step + 1] - E.line0;
          E.wg_off = wg;
          E.flags = ((r->outer == 0 && r->inner == 0) ? 1 : 0) | (j->prob.zero_skipping ? 2 : 0);
          wg += E.nlines;
          if (++r->step == r->nsteps) last.push_back(r);
          if (args.n == kTickMax && !flush()) bad = true;
        }
        if (!bad && !flush()) bad = true;
      }
      if (launched) {
        // the lane's run-ahead event, one every kTickEvery ticks (an event between
        // two launches delays the second)
        const int64_t t = lane_ticks[lane];
        if (t % kTickEvery == 0) (void)hipEventRecord(tick_ev[lane][(t / kTickEvery) % kTickEvents], stream);
        lane_ticks[lane]++;
        tick++;
        launched_any = true;
      }
    }
    if (bad) {
      fail_all(MBIR_ERR_HIP);
      continue;
    }
    if (launched_any) {
      std::lock_guard<std::mutex> lk(mu);
      st.ticks = tick;
      st.launches = n_launch;
      st.entries = n_entries;
      st.workgroups = n_wg;
    } else if (held) {
      // every lane with work is at its run-ahead bound: let the device catch up
      std::this_thread::yield();
    } else {
      // nothing to launch: wait for the oldest outstanding sweep end (of any lane)
      bool any_wait = false;
      for (Req* r : act)
        if (r->waiting || r->finishing) any_wait = true;
      if (any_wait) std::this_thread::yield();
    }
  }
  for (auto& l : tick_ev)
    for (auto& e : l) (void)hipEventDestroy(e);
}

// ---------------------------------------------------------------------------
extern "C" int64_t mbir_problem_device_bytes(const mbir_problem* p)
{
  if (!p) return MBIR_ERR_INPUT;
  return estimate_bytes(*p);
}

extern "C" int32_t mbir_server_create(mbir_ctx* ctx, const mbir_server_opts* o, mbir_server** out)
{
  if (!ctx || !out) return MBIR_ERR_INPUT;
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return MBIR_ERR_HIP;
  auto* s = new mbir_server();
  s->ctx = ctx;
  s->tune = ctx->tune;
  int threads = 2;
  if (o) {
    if (o->max_jobs > 0) s->max_jobs = std::min(o->max_jobs, kMaxServerJobs);
    if (o->setup_threads > 0) threads = std::min(o->setup_threads, 8);
    s->budget = o->hbm_budget_bytes;
    if (o->long_steps != 0) s->long_steps = o->long_steps;  // < 0: one lane
    if (o->long_jobs > 0) s->long_jobs = o->long_jobs;
  }
  if (s->budget == 0) s->budget = ctx->budget;
  if (s->budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return MBIR_ERR_HIP;
    }
    s->budget = (uint64_t)(0.8 * (double)fr);
  }
  // default-priority streams, one per lane
  if (hipStreamCreate(&s->lane_stream[0]) != hipSuccess ||
      hipStreamCreate(&s->lane_stream[1]) != hipSuccess ||
      s->djobs.alloc(s->max_jobs) != hipSuccess) {
    for (hipStream_t& q : s->lane_stream)
      if (q) (void)hipStreamDestroy(q);
    s->djobs.release();
    delete s;
    return MBIR_ERR_HIP;
  }
  for (int i = s->max_jobs - 1; i >= 0; --i) s->free_slots.push_back(i);
  s->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < threads; ++i) s->setup_thr.emplace_back([s] { s->setup_loop(); });
  s->run_thr = std::thread([s] { s->run_loop(); });
  *out = s;
  return MBIR_OK;
}

extern "C" int32_t mbir_server_submit(mbir_server* s, const mbir_problem* p, uint64_t* ticket)
{
  if (!s || !p || !ticket || !p->angles_deg || !p->counts) return MBIR_ERR_INPUT;
  if (p->bf_sinogram || p->estimate_prior) return MBIR_ERR_UNSUPPORTED;  // single-job paths only
  const mbir_geometry& g = p->geom;
  if (g.N_theta <= 0 || g.N_r <= 0 || g.N_t <= 0 || g.N_x <= 0 || g.N_y <= 0 || g.N_z <= 0)
    return MBIR_ERR_INPUT;
  auto* r = new Req();
  const size_t NT = g.N_theta, M = NT * g.N_r * g.N_t, V = (size_t)g.N_z * g.N_x * g.N_y;
  r->prob = *p;
  r->angles.assign(p->angles_deg, p->angles_deg + NT);
  r->counts.assign(p->counts, p->counts + M);
  r->prob.angles_deg = r->angles.data();
  r->prob.counts = r->counts.data();
  if (p->init_volume) {
    r->init.assign(p->init_volume, p->init_volume + V);
    r->prob.init_volume = r->init.data();
  }
  auto copy = [&](const double* src, std::vector<double>& dst, const double** field) {
    if (src) {
      dst.assign(src, src + NT);
      *field = dst.data();
    }
  };
  copy(p->init_gain, r->gain, &r->prob.init_gain);
  copy(p->init_offset, r->offset, &r->prob.init_offset);
  copy(p->init_variance, r->variance, &r->prob.init_variance);
  r->est = estimate_bytes(*p);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) {
    delete r;
    return MBIR_ERR_INPUT;
  }
  r->ticket = s->next_ticket++;
  r->t_submit = s->now();
  s->reqs[r->ticket] = r;
  s->queue.push_back(r);
  s->st.jobs_submitted++;
  *ticket = r->ticket;
  s->cv_setup.notify_all();
  return MBIR_OK;
}

static void fill_report(const Req* r, mbir_job_report* rep)
{
  if (!rep) return;
  std::memset(rep, 0, sizeof(*rep));
  rep->state = r->state;
  rep->error = r->error;
  rep->sweeps = r->sweeps;
  rep->outer_iterations = r->outer_done;
  rep->status = r->status;
  rep->voxel_visits = r->visits;
  rep->updated = r->updated;
  rep->device_bytes = r->job ? r->job->bytes : 0;
  rep->t_submit = r->t_submit;
  rep->t_admit = r->t_admit;
  rep->t_start = r->t_start;
  rep->t_done = r->t_done;
}

extern "C" int32_t mbir_server_wait(mbir_server* s, uint64_t ticket, double timeout_s,
                                    mbir_job_report* rep)
{
  if (!s) return MBIR_ERR_INPUT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MBIR_ERR_INPUT;
  Req* r = it->second;
  auto fin = [&] { return r->state == MBIR_JOB_DONE || r->state == MBIR_JOB_FAILED; };
  if (timeout_s < 0) s->cv_done.wait(lk, fin);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), fin)) {
    fill_report(r, rep);
    return 1;
  }
  fill_report(r, rep);
  return MBIR_OK;
}

extern "C" int32_t mbir_server_wait_any(mbir_server* s, double timeout_s, uint64_t* ticket)
{
  if (!s || !ticket) return MBIR_ERR_INPUT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto any = [&] { return !s->done_fifo.empty(); };
  if (timeout_s < 0) s->cv_done.wait(lk, any);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), any)) return 1;
  *ticket = s->done_fifo.front();
  s->done_fifo.pop_front();
  return MBIR_OK;
}

extern "C" int32_t mbir_server_job(mbir_server* s, uint64_t ticket, mbir_job** job)
{
  if (!s || !job) return MBIR_ERR_INPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second->state != MBIR_JOB_DONE) return MBIR_ERR_INPUT;
  *job = it->second->job;
  return MBIR_OK;
}

extern "C" int32_t mbir_server_release(mbir_server* s, uint64_t ticket)
{
  if (!s) return MBIR_ERR_INPUT;
  Req* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return MBIR_ERR_INPUT;
    r = it->second;
    if (r->state != MBIR_JOB_DONE && r->state != MBIR_JOB_FAILED) return MBIR_ERR_INPUT;
    s->reqs.erase(it);
  }
  if (r->job) mbir_job_destroy(r->job);
  if (r->hacc) (void)hipHostFree(r->hacc);
  if (r->hsched) (void)hipHostFree(r->hsched);
  if (r->ev) (void)hipEventDestroy(r->ev);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->reserved > 0) {
      s->reserved -= (uint64_t)r->reserved;
      s->resident--;
    }
    s->cv_setup.notify_all();
  }
  delete r;
  return MBIR_OK;
}

extern "C" int32_t mbir_server_cancel(mbir_server* s, uint64_t ticket)
{
  if (!s) return MBIR_ERR_INPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MBIR_ERR_INPUT;
  Req* r = it->second;
  if (r->state == MBIR_JOB_QUEUED) {
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), r), s->queue.end());
    s->finish(r, MBIR_JOB_FAILED, MBIR_ERR_CANCELLED);
  } else {
    r->cancel.store(true);
  }
  return MBIR_OK;
}

extern "C" int32_t mbir_server_get_stats(mbir_server* s, mbir_server_stats* out)
{
  if (!s || !out) return MBIR_ERR_INPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->reserved_bytes = s->reserved;
  out->peak_reserved_bytes = s->peak_reserved;
  out->budget_bytes = s->budget;
  out->fatal_error = s->fatal;
  return MBIR_OK;
}

extern "C" int32_t mbir_server_destroy(mbir_server* s)
{
  if (!s) return MBIR_ERR_INPUT;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    while (!s->queue.empty()) {
      Req* r = s->queue.front();
      s->queue.pop_front();
      s->finish(r, MBIR_JOB_FAILED, MBIR_ERR_CANCELLED);
    }
    for (auto& kv : s->reqs) kv.second->cancel.store(true);
    s->cv_setup.notify_all();
    s->cv_run.notify_all();
  }
  for (auto& t : s->setup_thr) t.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cv_run.notify_all();
  }
  s->run_thr.join();
  for (hipStream_t q : s->lane_stream) (void)hipStreamSynchronize(q);
  std::vector<uint64_t> tickets;
  for (auto& kv : s->reqs) tickets.push_back(kv.first);
  for (uint64_t t : tickets) mbir_server_release(s, t);
  s->djobs.release();
  for (hipStream_t q : s->lane_stream) (void)hipStreamDestroy(q);
  delete s;
  return MBIR_OK;
}
