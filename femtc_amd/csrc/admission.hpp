// Admission, queueing and eviction of the request server, as host code with
// no device dependency: the server drives it under its lock with real
// solvers and the real clock, and femtc_admission_simulate drives the same
// code with a fake clock and a fake allocator (CPU tests, sanitizer builds).
//
// Rules.  Requests are taken in arrival order, first admissible first: an
// apply is always admissible; a solve is admissible when its slot already
// holds a resident workspace for (problem, inner precision), or when its
// estimated workspace fits the budget after evicting idle workspaces least
// recently used first.  A solve that does not fit now waits while later
// requests that fit go ahead.  A workspace is charged at its estimate while
// it is being created, then at the size the caller reports.  Problems are
// known by keys the server issues; removing a problem retires its
// workspaces (idle ones at once, busy ones when their request ends), so a
// later problem can never meet a stale workspace.
#pragma once

#include <cstddef>
#include <cstdint>
#include <deque>
#include <vector>

namespace femtc
{

struct AdmReq
{
  int64_t id = 0;        // the caller's handle of the request
  int kind = 0;          // FEMTC_REQUEST_APPLY or FEMTC_REQUEST_SOLVE
  uint64_t problem = 0;  // key from add_problem
  int inner = 0;         // inner precision of a solve
  size_t bytes = 0;      // estimated workspace of a new solve
  double t_arrival = 0;
};

// What take() decided for a slot.
struct AdmTake
{
  AdmReq req;
  void* cached = nullptr; // resident workspace now busy, or null
  size_t reserved = 0;    // bytes reserved for a workspace to create
};

class Admission
{
public:
  explicit Admission(size_t budget) : budget_(budget) {}

  size_t budget() const { return budget_; }
  size_t used() const { return used_; }
  size_t high_water() const { return high_; }
  int resident() const { return (int)cache_.size(); }
  uint64_t evictions() const { return evictions_; }
  bool queue_empty() const { return queue_.empty(); }
  size_t queued() const { return queue_.size(); }

  uint64_t add_problem() { return ++next_key_; }
  bool known(uint64_t key) const;
  // Forget a problem: its idle workspaces go to victims (the caller
  // destroys them outside its lock), busy ones are destroyed when their
  // request ends, and its queued requests are dropped (their ids returned).
  void remove_problem(uint64_t key, std::vector<void*>* victims,
                      std::vector<int64_t>* dropped);

  void push(const AdmReq& r) { queue_.push_back(r); }
  // The first admissible queued request for `slot`, removed from the
  // queue; idle workspaces evicted to make room go to victims.
  bool take(int slot, AdmTake* out, std::vector<void*>* victims);
  // A request taken with a resident workspace ended; a workspace of a
  // removed problem is returned for destruction (else null).
  void* release(const AdmTake& t, int slot);
  // A workspace was created for a taken request: keep it resident, charged
  // at `actual` bytes; returned for destruction at once (non-null) if its
  // problem was removed meanwhile.
  void* adopt(const AdmTake& t, int slot, void* handle, size_t actual);
  // Creating the workspace failed: the reservation is returned.
  void cancel(const AdmTake& t);
  // Every workspace, for destruction (the server is closing).
  void drain(std::vector<void*>* victims);

private:
  struct Entry
  {
    uint64_t problem;
    int inner;
    int slot;
    void* handle;
    size_t bytes;
    uint64_t last_use;
    bool busy;
    bool retired; // problem removed while busy
  };
  int find(uint64_t problem, int inner, int slot) const;
  bool make_room(size_t need, std::vector<void*>* victims);
  void charge(size_t b);

  size_t budget_;
  size_t used_ = 0;
  size_t high_ = 0;
  uint64_t tick_ = 0;
  uint64_t next_key_ = 0;
  uint64_t evictions_ = 0;
  std::vector<uint64_t> removed_; // keys of removed problems
  std::deque<AdmReq> queue_;
  std::vector<Entry> cache_;
};

} // namespace femtc
