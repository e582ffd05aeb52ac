// queue_model.cpp -- a CPU model of the persistent kernels' work-queue protocol (csrc/device.h
// q_take / q_push / q_finish / claim / release_or_rerun, csrc/batch.hip k_visit_multi) with
// several persistent launches in flight on one device, for deadlock and lost-work checks.
//
// Each launch has its own queue (ring of slots, head = tickets | DONE, tail / pending), block
// flags and block states, and a solve: a random block graph whose "times" follow Bellman-Ford
// relaxations, so a visit lowers a block's value from its suppliers and queues its dependents
// when the value drops, as a real visit does.  A launch's grid fills the device (as the
// occupancy-sized grids do); the device has X XCDs with a fixed number of workgroup slots each,
// workgroup w of a launch goes to XCD w % X, and each XCD dispatches its pending workgroups in
// launch order then workgroup order whenever a slot frees (or, with -interleave, in a random
// order among launches).  A workgroup is a set of micro-threads whose every step is one atomic
// access, interleaved at random: thread 0 (take, visit, release), wave 0 (flag exchange and the
// push of dependents) and thread 64 (the next take issued while wave 0 pushes), with the
// workgroup barriers of the kernel between them.
//
// Checked: every launch ends (no state where all resident workgroups wait and nothing can be
// dispatched), every queued entry is consumed, pending returns to 0, and each solve reaches the
// Bellman-Ford fixed point of its graph.
//
// usage: queue_model <schedules> <launches> <seed> [-interleave]
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr uint32_t kDone = 0x80000000u;
constexpr int32_t kDoneMark = 0x7fffffff;
std::mt19937_64 rng;

int rnd(int n) { return (int)(rng() % (uint64_t)n); }

struct Graph {
  int nb = 0;
  std::vector<std::vector<int>> sup, dep;  // suppliers and dependents of each block
  std::vector<std::vector<int>> w;         // weight of each supplier edge
  std::vector<int> seeds;
};

Graph make_graph(int nb) {
  Graph g;
  g.nb = nb;
  g.sup.resize(nb);
  g.dep.resize(nb);
  g.w.resize(nb);
  // a ring plus random chords: every block depends on a few others, symmetric like halos
  for (int b = 0; b < nb; ++b) {
    int k = 1 + rnd(4);
    for (int j = 0; j < k; ++j) {
      int c = j == 0 ? (b + 1) % nb : rnd(nb);
      if (c == b) continue;
      if (std::find(g.dep[b].begin(), g.dep[b].end(), c) != g.dep[b].end()) continue;
      g.dep[b].push_back(c);
      g.dep[c].push_back(b);
      g.sup[c].push_back(b);
      g.sup[b].push_back(c);
    }
  }
  for (int b = 0; b < nb; ++b)
    for (size_t j = 0; j < g.sup[b].size(); ++j) g.w[b].push_back(1 + rnd(9));
  int ns = 1 + rnd(3);
  for (int i = 0; i < ns; ++i) g.seeds.push_back(rnd(nb));
  return g;
}

std::vector<int> fixed_point(const Graph& g) {
  std::vector<int> v(g.nb, 1 << 28);
  for (int s : g.seeds) v[s] = 0;
  for (bool ch = true; ch;) {
    ch = false;
    for (int b = 0; b < g.nb; ++b)
      for (size_t j = 0; j < g.sup[b].size(); ++j)
        if (v[g.sup[b][j]] + g.w[b][j] < v[b]) {
          v[b] = v[g.sup[b][j]] + g.w[b][j];
          ch = true;
        }
  }
  return v;
}

struct Launch {
  Graph g;
  uint32_t qmask = 0;
  std::vector<int32_t> slot;
  uint32_t head = 0;
  uint32_t tail = 0, pending = 0;  // one 64-bit word on the device
  std::vector<uint32_t> qflag, bstate;
  std::vector<int> val;  // the solve's "times"
  int grid = 0;
  int exited = 0;
  int dispatched = 0;
};

// the queue operations, one atomic step each (the device's agent-scope atomics are modelled as
// sequentially consistent)
void q_finish(Launch& L) {
  uint32_t h = L.head & ~kDone;
  L.head |= kDone;
  for (uint32_t i = L.tail; (int32_t)(h - i) > 0; ++i) L.slot[i & L.qmask] = kDoneMark;
}

void q_push_single(Launch& L, int b) {  // seeding
  if (L.qflag[b]) return;
  L.qflag[b] = 1;
  uint32_t t = L.tail++;
  ++L.pending;
  L.slot[t & L.qmask] = b + 1;
}

// Micro-thread states of a workgroup
enum Phase {
  kEntry, kTakeTicket, kTakeWait, kTakeClaim, kVisit, kFlag, kRelease, kBarrier1, kPush,
  kBarrier2, kExit
};

struct Take {  // an in-progress q_take
  int st = 0;  // 0 ticket, 1 wait, 2 claim
  uint32_t r = 0;
  int32_t v = 0;
  int result = -2;  // block, -1 finished
};

struct WG {
  int launch = 0, id = 0, xcd = 0;
  bool resident = false, done = false;
  int phase = kEntry;
  int b = -1;            // block under visit
  int next = -2;         // block taken by thread 64
  Take t0, t64;          // takes of thread 0 and thread 64
  bool converged = true, changed = false, again = false;
  std::vector<int> want; // dependents to push (after flag exchange)
  size_t push_i = 0;
  bool pushed_dec = false;
  bool t64_active = false, wave0_done = false;
};

struct Device {
  int X = 8, per_xcd = 2;
  std::vector<int> free_slots;
  std::vector<Launch> L;
  std::vector<WG> wgs;
  bool interleave = false;
};

// one step of a take; returns true when it finished (result set).  Steps: the ticket (atomic add
// on head), the wait on the ticket's slot and its clear, the clear of the block's queued flag,
// the claim (or the drop of an entry whose block is under visit).
bool take_step(Launch& L, Take& t) {
  switch (t.st) {
    case 0: {
      uint32_t r = L.head++;
      if (r & kDone) {
        t.result = -1;
        return true;
      }
      t.r = r;
      t.st = 1;
      return false;
    }
    case 1: {
      int32_t v = L.slot[t.r & L.qmask];
      if (v == 0) return false;  // still waiting
      L.slot[t.r & L.qmask] = 0;
      if (v == kDoneMark) {
        t.result = -1;
        return true;
      }
      t.v = v;
      t.st = 2;
      return false;
    }
    case 2:
      L.qflag[t.v - 1] = 0;
      t.st = 3;
      return false;
    default: {
      int b = t.v - 1;
      bool own = L.bstate[b]++ == 0;
      if (own) {
        t.result = b;
        return true;
      }
      uint32_t old = L.pending--;
      if (old == 1) q_finish(L);
      t.st = 0;
      return false;
    }
  }
}

bool take_blocked(const Launch& L, const Take& t) { return t.st == 1 && L.slot[t.r & L.qmask] == 0; }

// visit: Bellman-Ford relaxation of block b from its suppliers; converged at random (a sweep cap
// that cut the visit short re-queues the block itself)
void visit(Launch& L, WG& w) {
  const Graph& g = L.g;
  int b = w.b, best = L.val[b];
  for (size_t j = 0; j < g.sup[b].size(); ++j) best = std::min(best, L.val[g.sup[b][j]] + g.w[b][j]);
  w.changed = best < L.val[b];
  w.converged = true;
  if (w.changed && rnd(4) == 0) {
    // cut short: the value lands halfway, the block must run again
    best = L.val[b] - std::max(1, (L.val[b] - best) / 2);
    w.converged = false;
  }
  L.val[b] = best;
}

bool release_or_rerun(Launch& L, int b) {
  if (L.bstate[b] == 1) {
    L.bstate[b] = 0;
    return false;
  }
  L.bstate[b] -= L.bstate[b] - 1;
  return true;
}

// one random step of workgroup w; returns false if it could not move (waiting)
bool step(Device& D, WG& w) {
  Launch& L = D.L[w.launch];
  switch (w.phase) {
    case kEntry:
      if (L.pending == 0 && !(L.head & kDone)) q_finish(L);
      w.phase = kTakeTicket;
      w.t0 = Take();
      return true;
    case kTakeTicket: {  // thread 0 takes (or uses thread 64's block)
      if (w.next != -2) {
        w.t0.result = w.next;
        w.next = -2;
      } else {
        if (take_blocked(L, w.t0)) return false;
        if (!take_step(L, w.t0)) return true;
      }
      if (w.t0.result < 0) {
        w.phase = kExit;
        return true;
      }
      w.b = w.t0.result;
      w.phase = kVisit;
      return true;
    }
    case kVisit:
      visit(L, w);
      w.want.clear();
      w.push_i = 0;
      w.pushed_dec = false;
      w.phase = kFlag;
      return true;
    case kFlag:  // flags of the dependents (wave 0) while thread 0 releases (kRelease)
      if (w.changed)
        for (int c : L.g.dep[w.b])
          if (!L.qflag[c]) {
            L.qflag[c] = 1;
            w.want.push_back(c);
          }
      w.phase = kRelease;
      return true;
    case kRelease:
      w.again = release_or_rerun(L, w.b);
      if (!w.converged && !w.again) {
        if (!L.qflag[w.b]) {
          L.qflag[w.b] = 1;
          w.want.push_back(w.b);
        }
      }
      w.phase = kPush;
      w.t64_active = !w.again;
      w.t64 = Take();
      w.wave0_done = false;
      return true;
    case kPush: {
      // two micro-threads: wave 0 pushes (one atomic step per call), thread 64 takes
      bool can0 = !w.wave0_done;
      bool can64 = w.t64_active && !take_blocked(L, w.t64);
      if (!can0 && !can64) {
        if (w.t64_active) return false;  // thread 64 waits for a slot
        w.phase = kBarrier2;
        return true;
      }
      if (can0 && (!can64 || rnd(2))) {
        if (w.push_i < w.want.size()) {
          // one slot per step: the add of tail/pending then the store (modelled as one add for
          // the whole batch below, and stores one at a time)
          if (w.push_i == 0) {
            uint32_t n = (uint32_t)w.want.size();
            uint32_t dec = w.again ? 0u : 1u;
            uint32_t base = L.tail;
            L.tail += n;
            L.pending += n - dec;
            w.pushed_dec = true;
            // the stores follow, one per step, at base + i
            w.b = w.b;  // unchanged
            w.t0.r = base;
          }
          L.slot[(w.t0.r + (uint32_t)w.push_i) & L.qmask] = w.want[w.push_i] + 1;
          ++w.push_i;
          if (w.push_i == w.want.size()) w.wave0_done = true;
        } else {
          if (!w.again) {
            uint32_t old = L.pending--;
            if (old == 1) q_finish(L);
          }
          w.wave0_done = true;
        }
        return true;
      }
      if (take_step(L, w.t64)) {
        w.next = w.t64.result;
        w.t64_active = false;
      }
      return true;
    }
    case kBarrier2:
      if (w.again) {
        w.phase = kVisit;  // re-run of the same block
      } else {
        w.t0 = Take();
        w.phase = kTakeTicket;
      }
      return true;
    case kExit:
      w.done = true;
      w.resident = false;
      D.free_slots[w.xcd]++;
      ++L.exited;
      return true;
  }
  return false;
}

// dispatch: each XCD places its next pending workgroup (launch order, or random launch order)
void dispatch(Device& D) {
  for (int x = 0; x < D.X; ++x) {
    while (D.free_slots[x] > 0) {
      int pick = -1;
      std::vector<int> cand;
      for (size_t i = 0; i < D.wgs.size(); ++i) {
        WG& w = D.wgs[i];
        if (w.resident || w.done || w.xcd != x) continue;
        // in-order per launch on this XCD: only the lowest pending workgroup of each launch
        bool first = true;
        for (size_t j = 0; j < i; ++j)
          if (D.wgs[j].launch == w.launch && D.wgs[j].xcd == x && !D.wgs[j].resident && !D.wgs[j].done) first = false;
        if (!first) continue;
        cand.push_back((int)i);
        if (!D.interleave) break;  // launch order: the earliest launch's workgroup
      }
      if (cand.empty()) break;
      pick = cand[D.interleave ? rnd((int)cand.size()) : 0];
      D.wgs[pick].resident = true;
      D.free_slots[x]--;
    }
  }
}

int run_schedule(int nlaunch, bool interleave) {
  Device D;
  D.X = 2 + rnd(7);
  D.per_xcd = 1 + rnd(3);
  D.interleave = interleave;
  D.free_slots.assign(D.X, D.per_xcd);
  const int grid = D.X * D.per_xcd;  // each launch's grid fills the device
  D.L.resize(nlaunch);
  for (int l = 0; l < nlaunch; ++l) {
    Launch& L = D.L[l];
    L.g = make_graph(2 + rnd(24));
    uint32_t qcap = 1;
    while (qcap < 2u * ((uint32_t)L.g.nb + (uint32_t)grid) + 2u) qcap <<= 1;
    L.qmask = qcap - 1;
    L.slot.assign(qcap, 0);
    L.qflag.assign(L.g.nb, 0);
    L.bstate.assign(L.g.nb, 0);
    L.val.assign(L.g.nb, 1 << 28);
    L.grid = grid;
    // a launch may have no seeds of this kind (its kernel must still end)
    if (rnd(5) != 0)
      for (int s : L.g.seeds) {
        L.val[s] = 0;
        q_push_single(L, s);
        for (int c : L.g.dep[s]) q_push_single(L, c);
      }
    else
      L.g.seeds.clear();
    for (int w = 0; w < grid; ++w) {
      WG x;
      x.launch = l;
      x.id = w;
      x.xcd = w % D.X;
      D.wgs.push_back(x);
    }
  }
  for (long steps = 0;; ++steps) {
    dispatch(D);
    std::vector<int> res;
    for (size_t i = 0; i < D.wgs.size(); ++i)
      if (D.wgs[i].resident) res.push_back((int)i);
    bool all_done = true;
    for (const WG& w : D.wgs) all_done &= w.done;
    if (all_done) break;
    if (res.empty()) {
      std::printf("DEADLOCK: workgroups pending but none resident\n");
      return 1;
    }
    // try resident workgroups in random order until one moves
    std::shuffle(res.begin(), res.end(), rng);
    bool moved = false;
    for (int i : res)
      if (step(D, D.wgs[i])) {
        moved = true;
        break;
      }
    if (!moved) {
      std::printf("DEADLOCK: %zu resident workgroups all wait\n", res.size());
      for (int l = 0; l < nlaunch; ++l)
        std::printf("  launch %d: head %u%s tail %u pending %u exited %d/%d\n", l,
                    D.L[l].head & ~kDone, D.L[l].head & kDone ? "|DONE" : "", D.L[l].tail,
                    D.L[l].pending, D.L[l].exited, D.L[l].grid);
      return 1;
    }
    if (steps > 50000000) {
      std::printf("LIVELOCK: no end after %ld steps\n", steps);
      return 1;
    }
  }
  for (int l = 0; l < nlaunch; ++l) {
    Launch& L = D.L[l];
    if (L.pending != 0) {
      std::printf("LOST: launch %d ended with pending %u\n", l, L.pending);
      return 1;
    }
    std::vector<int> ref = fixed_point(L.g);
    if (ref != L.val) {
      std::printf("WRONG: launch %d did not reach the fixed point\n", l);
      return 1;
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 2000;
  const int nl = argc > 2 ? std::atoi(argv[2]) : 4;
  const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1;
  const bool inter = argc > 4 && std::strcmp(argv[4], "-interleave") == 0;
  rng.seed(seed);
  for (int i = 0; i < n; ++i) {
    if (run_schedule(1 + rnd(nl), inter)) {
      std::printf("schedule %d failed\n", i);
      return 2;
    }
  }
  std::printf("%d schedules, up to %d launches in flight%s: every launch ended, no lost entry, "
              "every solve at its fixed point\n", n, nl, inter ? " (interleaved dispatch)" : "");
  return 0;
}
