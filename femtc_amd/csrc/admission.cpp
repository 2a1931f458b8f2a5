// Admission core of the request server (admission.hpp) and its fake-clock,
// fake-allocator simulator femtc_admission_simulate.  Host code only: it
// includes no HIP header and allocates no device memory, so it also builds
// with plain g++ and the sanitizers.
#include "admission.hpp"

#include <algorithm>
#include <cstdint>
#include <limits>
#include <vector>

#include "femtc.h"

namespace femtc
{

bool Admission::known(uint64_t key) const
{
  return key >= 1 && key <= next_key_ &&
         std::find(removed_.begin(), removed_.end(), key) == removed_.end();
}

int Admission::find(uint64_t problem, int inner, int slot) const
{
  for (size_t i = 0; i < cache_.size(); ++i)
    if (cache_[i].problem == problem && cache_[i].inner == inner &&
        cache_[i].slot == slot && !cache_[i].retired)
      return (int)i;
  return -1;
}

void Admission::charge(size_t b)
{
  used_ += b;
  high_ = std::max(high_, used_);
}

bool Admission::make_room(size_t need, std::vector<void*>* victims)
{
  while (used_ + need > budget_)
  {
    int victim = -1;
    for (size_t i = 0; i < cache_.size(); ++i)
      if (!cache_[i].busy &&
          (victim < 0 || cache_[i].last_use < cache_[victim].last_use))
        victim = (int)i;
    if (victim < 0)
      return false;
    victims->push_back(cache_[victim].handle);
    used_ -= cache_[victim].bytes;
    cache_.erase(cache_.begin() + victim);
    ++evictions_;
  }
  return true;
}

bool Admission::take(int slot, AdmTake* out, std::vector<void*>* victims)
{
  for (auto it = queue_.begin(); it != queue_.end(); ++it)
  {
    AdmTake t;
    t.req = *it;
    if (it->kind == FEMTC_REQUEST_SOLVE)
    {
      const int c = find(it->problem, it->inner, slot);
      if (c >= 0)
      {
        if (cache_[c].busy)
          continue; // cannot happen within one slot; kept safe
        cache_[c].busy = true;
        cache_[c].last_use = ++tick_;
        t.cached = cache_[c].handle;
      }
      else
      {
        if (it->bytes > budget_ || !make_room(it->bytes, victims))
          continue; // does not fit now: try a later request
        charge(it->bytes);
        t.reserved = it->bytes;
      }
    }
    queue_.erase(it);
    *out = t;
    return true;
  }
  return false;
}

void* Admission::release(const AdmTake& t, int slot)
{
  for (size_t i = 0; i < cache_.size(); ++i)
    if (cache_[i].handle == t.cached && cache_[i].slot == slot)
    {
      if (cache_[i].retired)
      {
        void* h = cache_[i].handle;
        used_ -= cache_[i].bytes;
        cache_.erase(cache_.begin() + i);
        return h;
      }
      cache_[i].busy = false;
      return nullptr;
    }
  return nullptr;
}

void* Admission::adopt(const AdmTake& t, int slot, void* handle,
                       size_t actual)
{
  used_ -= t.reserved;
  if (!known(t.req.problem))
    return handle; // removed while it was being created
  charge(actual);
  cache_.push_back(Entry{t.req.problem, t.req.inner, slot, handle, actual,
                         ++tick_, false, false});
  return nullptr;
}

void Admission::cancel(const AdmTake& t) { used_ -= t.reserved; }

void Admission::remove_problem(uint64_t key, std::vector<void*>* victims,
                               std::vector<int64_t>* dropped)
{
  if (!known(key))
    return;
  removed_.push_back(key);
  for (size_t i = 0; i < cache_.size();)
  {
    if (cache_[i].problem != key)
    {
      ++i;
      continue;
    }
    if (cache_[i].busy)
    {
      cache_[i].retired = true;
      ++i;
      continue;
    }
    victims->push_back(cache_[i].handle);
    used_ -= cache_[i].bytes;
    cache_.erase(cache_.begin() + i);
  }
  for (auto it = queue_.begin(); it != queue_.end();)
  {
    if (it->problem == key)
    {
      dropped->push_back(it->id);
      it = queue_.erase(it);
    }
    else
      ++it;
  }
}

void Admission::drain(std::vector<void*>* victims)
{
  for (const Entry& e : cache_)
    victims->push_back(e.handle);
  cache_.clear();
  used_ = 0;
}

} // namespace femtc

using namespace femtc;

extern "C" int femtc_admission_simulate(const femtc_sim_request* reqs, int n,
                                        int nproblems,
                                        const double* remove_at, int slots,
                                        size_t budget, femtc_sim_result* out,
                                        femtc_sim_stats* stats)
{
  if (n < 0 || nproblems < 0 || slots < 1 || (n > 0 && (!reqs || !out)))
    return FEMTC_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n; ++i)
    if (reqs[i].problem < 0 || reqs[i].problem >= nproblems ||
        (reqs[i].kind != FEMTC_REQUEST_APPLY &&
         reqs[i].kind != FEMTC_REQUEST_SOLVE) ||
        !(reqs[i].t_arrival >= 0.0) || !(reqs[i].service >= 0.0) ||
        !(reqs[i].create >= 0.0))
      return FEMTC_ERR_INVALID_ARGUMENT;
  const double inf = std::numeric_limits<double>::infinity();
  Admission adm(budget);
  std::vector<uint64_t> key(nproblems);
  for (int p = 0; p < nproblems; ++p)
    key[p] = adm.add_problem();
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return reqs[a].t_arrival < reqs[b].t_arrival;
  });
  for (int i = 0; i < n; ++i)
    out[i] = femtc_sim_result{FEMTC_ERR_NOT_CONVERGED, -1, 0, -1.0, -1.0};
  struct Run
  {
    bool busy = false;
    AdmTake t;
    double t_end = 0;
  };
  std::vector<Run> run(slots);
  std::vector<bool> removed(nproblems, false);
  std::vector<void*> victims;
  std::vector<int64_t> dropped;
  int64_t created = 0, destroyed = 0;
  size_t next = 0; // next arrival in `order`
  double now = 0.0;
  for (;;)
  {
    double t = next < order.size() ? reqs[order[next]].t_arrival : inf;
    for (const Run& r : run)
      if (r.busy)
        t = std::min(t, r.t_end);
    for (int p = 0; p < nproblems; ++p)
      if (remove_at && !removed[p] && remove_at[p] >= 0.0)
        t = std::min(t, remove_at[p]);
    if (t == inf)
      break;
    now = t;
    // completions
    for (int s = 0; s < slots; ++s)
    {
      Run& r = run[s];
      if (!r.busy || r.t_end > now)
        continue;
      const femtc_sim_request& q = reqs[r.t.req.id];
      void* dead = nullptr;
      if (q.kind == FEMTC_REQUEST_SOLVE)
      {
        if (r.t.cached)
          dead = adm.release(r.t, s);
        else
        {
          void* h = reinterpret_cast<void*>((uintptr_t)(++created));
          dead = adm.adopt(r.t, s, h, q.actual);
        }
      }
      if (dead)
        ++destroyed;
      r.busy = false;
    }
    // problem removals
    for (int p = 0; p < nproblems; ++p)
      if (remove_at && !removed[p] && remove_at[p] >= 0.0 &&
          remove_at[p] <= now)
      {
        removed[p] = true;
        adm.remove_problem(key[p], &victims, &dropped);
      }
    for (int64_t id : dropped)
      out[id].status = FEMTC_ERR_INVALID_ARGUMENT;
    dropped.clear();
    // arrivals
    while (next < order.size() && reqs[order[next]].t_arrival <= now)
    {
      const int i = order[next++];
      const femtc_sim_request& q = reqs[i];
      if (removed[q.problem])
        out[i].status = FEMTC_ERR_INVALID_ARGUMENT;
      else if (q.kind == FEMTC_REQUEST_SOLVE && q.bytes > budget)
        out[i].status = FEMTC_ERR_OUT_OF_MEMORY; // refused at submit
      else
        adm.push(AdmReq{i, q.kind, key[q.problem], q.inner, q.bytes,
                        q.t_arrival});
    }
    // free slots take the first admissible requests
    for (int s = 0; s < slots; ++s)
    {
      Run& r = run[s];
      if (r.busy || !adm.take(s, &r.t, &victims))
        continue;
      const int i = (int)r.t.req.id;
      const femtc_sim_request& q = reqs[i];
      const bool fresh = q.kind == FEMTC_REQUEST_SOLVE && !r.t.cached;
      r.busy = true;
      r.t_end = now + q.service + (fresh ? q.create : 0.0);
      out[i] = femtc_sim_result{FEMTC_OK, s, r.t.cached ? 1 : 0, now,
                                r.t_end};
    }
    destroyed += (int64_t)victims.size();
    victims.clear();
  }
  if (stats)
  {
    stats->high_water = adm.high_water();
    stats->evictions = (int64_t)adm.evictions();
    stats->created = created;
    stats->destroyed = destroyed;
    stats->makespan = now;
    stats->left_in_queue = (int64_t)adm.queued();
  }
  return FEMTC_OK;
}
