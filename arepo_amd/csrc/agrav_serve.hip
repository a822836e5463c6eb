// agrav_serve.hip -- continuous-batching server for many independent gravity
// solves (host runtime, no device code of its own).
//
// Requests (one particle set each) are queued in arrival order.  A fixed number
// of slots each own a resident solver context -- its own HIP stream and its own
// device buffers, which stay allocated between requests -- and a worker thread.
// Admission bounds the HBM the slots hold.  Each slot is charged what its
// context actually holds (the sum of its device buffers, measured after every
// solve), or, while it solves, what the request's footprint model says the
// buffers grow to.  A free slot admits the oldest queued request that fits the
// budget (first fit in arrival order); when only the buffers idle slots keep
// resident stand in the way, those contexts are replaced by fresh ones.  So
// requests that arrive while others run start as soon as a slot and memory free
// up, and their device work overlaps on the GPU through the slots' streams.
// Every solve runs the same deterministic kernels on its own buffers with the
// same softening table, so a result does not depend on what else is in flight
// or on what the slot solved before.
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
  size_t footprint = 0;  // device bytes of a context that has solved this request (model)
  agrav_status status = AGRAV_OK;
  bool done           = false;
  int slot            = -1;
};

}  // namespace

struct agrav_server
{
  agrav_params par{};
  size_t budget = 0, peak = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job *> queue;
  std::unordered_map<uint64_t, Job *> jobs;
  std::vector<std::thread> workers;
  std::vector<agrav_ctx *> slots;
  std::vector<size_t> resident;  // device bytes each slot's context holds (measured after every solve)
  std::vector<Job *> busy;       // the job each slot runs, or null
  size_t base = 0;               // device bytes of a fresh context (tables, PM meshes)
  uint64_t next_ticket = 1;
  bool stop            = false;
  long long completed = 0, failed = 0, evictions = 0, model_exceeded = 0;
  int running = 0, peak_running = 0;
  std::string err;
};

namespace
{

// Bytes slot s holds at most once job j has run on it.  Buffers only grow while a
// context lives: a fresh context ends within j's footprint; a context grown by
// earlier jobs keeps what it holds, and the buffers sized by other quantities than the
// particle count (groups, chunks, tie runs) may still grow past it, by far less than
// 16 B per particle of j.
size_t job_bound(const agrav_server *S, int s, const Job *j)
{
  const size_t res = S->resident[s];
  size_t b         = res > j->footprint ? res : j->footprint;
  if(res > S->base)
    b += (size_t)j->req.n * 16 + ((size_t)1 << 20);
  return b;
}

// What a slot is charged: its resident buffers, or the bound of its running job.
size_t charge(const agrav_server *S, int s)
{
  return S->busy[s] ? job_bound(S, s, S->busy[s]) : S->resident[s];
}

size_t in_use(const agrav_server *S)
{
  size_t t = 0;
  for(size_t s = 0; s < S->slots.size(); s++)
    t += charge(S, (int)s);
  return t;
}

// Drop an idle slot's resident buffers: the context is replaced by a fresh one
// (its worker is waiting on the server mutex, which the caller holds).
bool evict(agrav_server *S, int s)
{
  agrav_ctx *c = S->slots[s];
  if(c)
    agrav_destroy(c);
  S->slots[s]    = nullptr;
  S->resident[s] = 0;
  agrav_ctx *f   = nullptr;
  if(agrav_create(&S->par, &f) != AGRAV_OK)
    return false;
  S->slots[s]    = f;
  S->resident[s] = agrav::device_bytes(f);
  S->evictions++;
  return true;
}

bool fits(const agrav_server *S, int s, const Job *j)
{
  const size_t mine = S->resident[s];  // slot s is idle: charged its resident bytes
  return in_use(S) - mine + job_bound(S, s, j) <= S->budget;
}

// First queued job (arrival order) that slot s can admit within the budget,
// evicting idle slots' buffers (its own first, then the largest) when that makes
// it fit.  Null when every queued job must wait for a running job to end.
Job *pick_job(agrav_server *S, int s)
{
  if(!S->slots[s] && !evict(S, s))
    return nullptr;
  for(auto it = S->queue.begin(); it != S->queue.end(); ++it)
    {
      Job *j = *it;
      if(!fits(S, s, j))
        {
          // the least the server can hold with j on slot s: every other idle
          // slot back to a fresh context, the running ones as charged
          size_t floor = j->footprint;
          for(size_t o = 0; o < S->slots.size(); o++)
            if((int)o != s)
              floor += S->busy[o] ? charge(S, (int)o) : S->base;
          if(floor > S->budget)
            continue;  // fits only after a running job ends
          if(S->resident[s] > j->footprint && !evict(S, s))
            return nullptr;
          while(!fits(S, s, j))
            {
              int victim = -1;
              for(size_t o = 0; o < S->slots.size(); o++)
                if((int)o != s && !S->busy[o] && S->resident[o] > S->base &&
                   (victim < 0 || S->resident[o] > S->resident[victim]))
                  victim = (int)o;
              if(victim < 0 || !evict(S, victim))
                return nullptr;
            }
        }
      S->queue.erase(it);
      return j;
    }
  return nullptr;
}

void run_job(agrav_server *S, agrav_ctx *c, Job *j)
{
  const agrav_request &r = j->req;
  agrav_status s         = c ? AGRAV_OK : AGRAV_ERR_STATE;
  // the effective softening table: the server's, with type 0 replaced by the
  // request's value when it gives one; set whenever the slot's table differs, so
  // a result never depends on what the slot solved before
  double eps[AGRAV_MAX_SOFT_TYPES];
  const int ns = S->par.n_soft_types;
  for(int k = 0; k < ns; k++)
    eps[k] = S->par.softening[k];
  if(r.softening > 0)
    eps[0] = r.softening;
  bool same = c && c->par.n_soft_types == ns;
  for(int k = 0; same && k < ns; k++)
    same = c->par.softening[k] == eps[k];
  if(s == AGRAV_OK && !same)
    s = agrav_set_softenings(c, ns, eps);
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
  std::unique_lock<std::mutex> lk(S->mu);
  for(;;)
    {
      Job *pick = nullptr;
      for(;;)
        {
          if(S->queue.empty())
            {
              if(S->stop)
                return;
              S->cv_work.wait(lk);
              continue;
            }
          if((pick = pick_job(S, slot)))
            break;
          if(S->running == 0)
            {
              // nothing will free memory (a fresh context could not be made):
              // the oldest request fails instead of waiting for ever
              Job *j = S->queue.front();
              S->queue.pop_front();
              j->status = AGRAV_ERR_OOM;
              j->t_start = j->t_end = clk::now();
              j->done               = true;
              S->failed++;
              S->cv_done.notify_all();
              continue;
            }
          // every queued job waits for memory that a running job holds; its end
          // notifies cv_work (with nothing running, every job fits: submit
          // rejects footprints above the budget)
          S->cv_work.wait(lk);
        }
      S->busy[slot] = pick;
      S->running++;
      S->peak_running = S->running > S->peak_running ? S->running : S->peak_running;
      const size_t u  = in_use(S);
      S->peak         = u > S->peak ? u : S->peak;
      pick->slot      = slot;
      pick->t_start   = clk::now();
      agrav_ctx *c    = S->slots[slot];
      lk.unlock();
      run_job(S, c, pick);
      const size_t now = c ? agrav::device_bytes(c) : 0;
      lk.lock();
      pick->t_end = clk::now();
      pick->done  = true;
      // the bound charged while the job ran; count the solves that exceed it (their
      // real bytes are charged from now on)
      const size_t bound = job_bound(S, slot, pick);
      if(now > bound)
        S->model_exceeded++;
      S->resident[slot] = now;
      S->busy[slot]     = nullptr;
      const size_t u2   = in_use(S);
      S->peak           = u2 > S->peak ? u2 : S->peak;
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

namespace agrav
{
// Device bytes one solve of n particles adds to a fresh context: about 350 B per
// particle of inputs, sorted copies, keys, prefix sums, results and the walk's
// records at child-list positions, about 225 B per octree node (at most 1.5 nodes
// per particle assumed; clustered sets have 0.5-0.7), the sort's scratch and a fixed
// reserve for the small buffers.
size_t solve_footprint(const agrav_params *p, int64_t n)
{
  const size_t nn = (size_t)(n > 0 ? n : 1);
  size_t per      = 860;
  if(p && p->pmgrid > 0)
    per += 32;  // PM accelerations and potentials per particle
  return nn * per + ((size_t)16 << 20);
}
}  // namespace agrav

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
      S->resident.push_back(agrav::device_bytes(c));
      S->busy.push_back(nullptr);
    }
  S->base = S->resident[0];
  if((size_t)slots * S->base > S->budget)
    {
      for(agrav_ctx *d : S->slots)
        agrav_destroy(d);
      delete S;
      agrav::set_error_msg("the HBM budget does not hold the slots' fresh contexts");
      return AGRAV_ERR_CAPACITY;
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
  j->footprint = S->base + agrav::solve_footprint(&S->par, r->n);
  j->t_submit  = clk::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if(S->stop)
      {
        delete j;
        return AGRAV_ERR_STATE;
      }
    // with every other slot at a fresh context, the request must fit the budget
    if(j->footprint + (S->slots.size() - 1) * S->base > S->budget)
      {
        delete j;
        agrav::set_error_msg("the request's HBM footprint exceeds the server's budget");
        return AGRAV_ERR_CAPACITY;
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
  st->hbm_in_use      = in_use(S);
  st->hbm_peak        = S->peak;
  st->hbm_budget      = S->budget;
  st->slots           = (int)S->slots.size();
  size_t res          = 0;
  for(size_t s = 0; s < S->slots.size(); s++)
    res += S->resident[s];
  st->hbm_resident    = res;
  st->evictions       = S->evictions;
  st->model_exceeded  = S->model_exceeded;
  return AGRAV_OK;
}

size_t agrav_solve_footprint(const agrav_params *p, int64_t n) { return agrav::solve_footprint(p, n); }

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
    if(c)
      agrav_destroy(c);
  for(auto &kv : S->jobs)
    delete kv.second;
  delete S;
}

}  // extern "C"
