// agrav_serve.hip -- continuous-batching server for many independent gravity
// solves (host runtime, no device code of its own).
//
// Requests (one particle set each) are queued in arrival order.  A fixed number
// of slots each own a resident solver context -- its own HIP stream and its own
// device buffers, which stay allocated between requests -- and a worker thread.
// A free slot admits the oldest queued request whose estimated HBM footprint fits
// the remaining budget (first fit in arrival order), so requests that arrive
// while others run start as soon as a slot and memory free up, and their device
// work overlaps on the GPU through the slots' streams.  Every solve runs the same
// deterministic kernels on its own buffers, so a result does not depend on what
// else is in flight.
//
// The reference solves independent sets one after another on one core
// (subfind_unbind, subfind_serial.c:429-560); this is the GPU serving analogue.
#include "agrav_ctx.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace
{

using clk = std::chrono::steady_clock;

struct Job
{
  uint64_t ticket;
  agrav_request req;
  clk::time_point t_submit, t_start, t_end;
  size_t footprint = 0;
  agrav_status status = AGRAV_OK;
  bool done           = false;
  int slot            = -1;
};

}  // namespace

struct agrav_server
{
  agrav_params par{};
  size_t budget = 0, in_use = 0, peak = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job *> queue;
  std::unordered_map<uint64_t, Job *> jobs;
  std::vector<std::thread> workers;
  std::vector<agrav_ctx *> slots;
  uint64_t next_ticket = 1;
  bool stop            = false;
  long long completed = 0, failed = 0;
  int running = 0, peak_running = 0;
  std::string err;
};

namespace
{

// HBM footprint of one solve: particles (inputs, sorted copies, keys, results),
// octree (about 0.5 nodes per particle for clustered sets, 1 worst case) and scratch
size_t footprint_of(int64_t n) { return (size_t)(n > 0 ? n : 1) * 720 + ((size_t)64 << 20); }

void run_job(agrav_ctx *c, Job *j)
{
  const agrav_request &r = j->req;
  agrav_status s         = AGRAV_OK;
  if(r.softening > 0 && r.softening != c->par.softening[0])
    s = agrav_set_softenings(c, 1, &r.softening);
  if(s == AGRAV_OK)
    s = agrav_set_particles(c, r.n, r.pos, r.mass, r.softtype);
  if(s == AGRAV_OK)
    s = agrav_compute_grav_accelerations(c, 1);
  if(s == AGRAV_OK)
    s = agrav_get_accelerations(c, r.acc, r.pot, nullptr, nullptr);
  j->status = s;
}

void worker(agrav_server *S, int slot)
{
  agrav_ctx *c = S->slots[slot];
  std::unique_lock<std::mutex> lk(S->mu);
  for(;;)
    {
      Job *pick = nullptr;
      S->cv_work.wait(lk, [&] {
        if(S->stop)
          return true;
        for(Job *j : S->queue)
          if(S->in_use + j->footprint <= S->budget || S->running == 0)
            return true;
        return false;
      });
      if(S->stop && S->queue.empty())
        return;
      for(auto it = S->queue.begin(); it != S->queue.end(); ++it)
        if(S->in_use + (*it)->footprint <= S->budget || S->running == 0)
          {
            pick = *it;
            S->queue.erase(it);
            break;
          }
      if(!pick)
        continue;
      S->in_use += pick->footprint;
      S->peak = S->in_use > S->peak ? S->in_use : S->peak;
      S->running++;
      S->peak_running = S->running > S->peak_running ? S->running : S->peak_running;
      pick->slot    = slot;
      pick->t_start = clk::now();
      lk.unlock();
      run_job(c, pick);
      lk.lock();
      pick->t_end = clk::now();
      pick->done  = true;
      S->in_use -= pick->footprint;
      S->running--;
      if(pick->status == AGRAV_OK)
        S->completed++;
      else
        S->failed++;
      S->cv_done.notify_all();
      S->cv_work.notify_all();
    }
}

double ms_between(clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

}  // namespace

extern "C" {

agrav_status agrav_server_create(const agrav_params *p, int slots, size_t hbm_budget_bytes, agrav_server **out)
{
  if(!p || !out || slots < 1 || slots > 64)
    return AGRAV_ERR_BAD_PARAM;
  *out            = nullptr;
  agrav_server *S = new agrav_server();
  S->par          = *p;
  S->budget       = hbm_budget_bytes ? hbm_budget_bytes : ((size_t)64 << 30);
  for(int i = 0; i < slots; i++)
    {
      agrav_ctx *c    = nullptr;
      agrav_status st = agrav_create(p, &c);
      if(st != AGRAV_OK)
        {
          for(agrav_ctx *d : S->slots)
            agrav_destroy(d);
          delete S;
          return st;
        }
      S->slots.push_back(c);
    }
  for(int i = 0; i < slots; i++)
    S->workers.emplace_back(worker, S, i);
  *out = S;
  return AGRAV_OK;
}

agrav_status agrav_server_submit(agrav_server *S, const agrav_request *r, uint64_t *ticket)
{
  if(!S || !r || !ticket || r->n < 0 || (r->n > 0 && (!r->pos || !r->mass || !r->acc)))
    return AGRAV_ERR_BAD_PARAM;
  Job *j       = new Job();
  j->req       = *r;
  j->footprint = footprint_of(r->n);
  j->t_submit  = clk::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if(S->stop)
      {
        delete j;
        return AGRAV_ERR_STATE;
      }
    j->ticket = S->next_ticket++;
    S->jobs[j->ticket] = j;
    S->queue.push_back(j);
    *ticket = j->ticket;
  }
  S->cv_work.notify_one();
  return AGRAV_OK;
}

static void fill_result(const Job *j, agrav_request_result *res)
{
  res->status   = j->status;
  res->queue_ms = ms_between(j->t_submit, j->t_start);
  res->run_ms   = ms_between(j->t_start, j->t_end);
  res->total_ms = ms_between(j->t_submit, j->t_end);
  res->slot     = j->slot;
}

agrav_status agrav_server_wait(agrav_server *S, uint64_t ticket, agrav_request_result *res)
{
  if(!S)
    return AGRAV_ERR_BAD_PARAM;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(ticket);
  if(it == S->jobs.end())
    return AGRAV_ERR_BAD_PARAM;
  Job *j = it->second;
  S->cv_done.wait(lk, [&] { return j->done; });
  if(res)
    fill_result(j, res);
  agrav_status s = j->status;
  S->jobs.erase(it);
  delete j;
  return s;
}

agrav_status agrav_server_poll(agrav_server *S, uint64_t ticket, int *done)
{
  if(!S || !done)
    return AGRAV_ERR_BAD_PARAM;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->jobs.find(ticket);
  if(it == S->jobs.end())
    return AGRAV_ERR_BAD_PARAM;
  *done = it->second->done ? 1 : 0;
  return AGRAV_OK;
}

agrav_status agrav_server_stats(agrav_server *S, agrav_server_info *st)
{
  if(!S || !st)
    return AGRAV_ERR_BAD_PARAM;
  std::lock_guard<std::mutex> lk(S->mu);
  st->completed       = S->completed;
  st->failed          = S->failed;
  st->queued          = (int)S->queue.size();
  st->running         = S->running;
  st->peak_running    = S->peak_running;
  st->hbm_in_use      = S->in_use;
  st->hbm_peak        = S->peak;
  st->hbm_budget      = S->budget;
  st->slots           = (int)S->slots.size();
  return AGRAV_OK;
}

void agrav_server_destroy(agrav_server *S)
{
  if(!S)
    return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for(auto &t : S->workers)
    t.join();
  for(agrav_ctx *c : S->slots)
    agrav_destroy(c);
  for(auto &kv : S->jobs)
    delete kv.second;
  delete S;
}

}  // extern "C"
