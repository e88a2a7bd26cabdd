// CPU check of the batched closure's schedule: the library's plan builder
// (fapi::build_fw_batch) and the task -> tile maps the kernels use
// (batch_desc.h) drive a host simulation of every launch on small random
// matrices.  Checked per launch: the tasks write distinct tiles, no task
// reads a tile another task of the launch writes, and every segment runs on
// the launch's one kernel instance; at the end every matrix equals plain
// Floyd-Warshall (fap/kernel/floyd.h:38-49, k-i-j order) bit for bit.
//
// Second part: a model of ONE persistent launch running the same tasks as
// the library queues them for fw_queue_kernel (fapi::build_fw_queue: every
// job's phases, launch after launch, each segment with the number of its job's tasks that must complete first,
// checked against the launch sequence), in which
// a workgroup waits (spinning on its job's completion counter) for its
// job's earlier phases.  The hardware dispatches a grid's workgroups in order of
// their id as slots free up and never preempts a resident one, so at most
// R workgroups make progress at a time.  The model replays exactly that
// (R slots, in-order dispatch, one task step per running workgroup per
// tick) and reports whether the launch drains:
//  * tasks dealt statically (workgroup b takes tasks b, b + G, ... of the
//    phase-ordered list) with G > R workgroups: resident workgroups can all
//    wait for tasks held by workgroups that are not yet dispatched, which
//    can only start when a resident one exits -- the launch never ends.
//    The model must find that state (it is how the removed design hung);
//  * the same with G <= R (every workgroup resident): drains;
//  * tasks claimed in list order from a shared counter (a workgroup claims
//    its next task only when it is running): drains for any G and R, since
//    the first unfinished task is always claimed, by a running workgroup,
//    and everything it waits for comes earlier in the list; the same when,
//    as in the kernel, the next task is claimed once the current task's
//    wait is over (the holder of the first unfinished task is then running
//    it, or has finished everything before it);
//  * every wait bounded by a spin limit that, when exceeded, raises an
//    error flag and exits the workgroup: the launch always ends, and the
//    host sees the flag.
// Tasks run on the live matrices in the order the model completes them;
// where the launch drains, every matrix must equal plain Floyd-Warshall.
// Prints "ok" and exits 0 when everything holds.
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <random>
#include <set>
#include <utility>
#include <vector>

#include "batch_desc.h"

using fapk::BT;
static const int32_t kInf = 0x3FFFFFFF;

struct Mat {
  int n = 0;
  std::vector<int32_t> a;
  int32_t get(int i, int j) const { return a[(size_t)i * n + j]; }
  int32_t &at(int i, int j) { return a[(size_t)i * n + j]; }
};

static int lo(int b) { return b * BT; }
static int hi(const Mat &m, int b) { return std::min(m.n, (b + 1) * BT); }

// Tile (I, J) of D = min(S, S[tile rows][k] + S[k][tile cols]) over the
// pivot blocks [p0, p1], operands from the launch-start snapshot S.
static void tile_update(const Mat &S, Mat &D, int I, int J, int p0, int p1) {
  for (int i = lo(I); i < hi(S, I); i++)
    for (int j = lo(J); j < hi(S, J); j++) {
      int32_t v = S.get(i, j);
      for (int k = lo(p0); k < hi(S, p1); k++) v = std::min(v, S.get(i, k) + S.get(k, j));
      D.at(i, j) = v;
    }
}

static void close_diag(Mat &D, int r) {
  for (int k = lo(r); k < hi(D, r); k++)
    for (int i = lo(r); i < hi(D, r); i++)
      for (int j = lo(r); j < hi(D, r); j++) D.at(i, j) = std::min(D.at(i, j), D.get(i, k) + D.get(k, j));
}

static Mat plain_fw(Mat m) {
  for (int k = 0; k < m.n; k++)
    for (int i = 0; i < m.n; i++)
      for (int j = 0; j < m.n; j++) m.at(i, j) = std::min(m.at(i, j), m.get(i, k) + m.get(k, j));
  return m;
}

static Mat random_matrix(int n, uint32_t seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::uniform_int_distribution<int> w(1, 1000);
  Mat m;
  m.n = n;
  m.a.assign((size_t)n * n, kInf);
  const double p = std::min(1.0, 6.0 / std::max(n, 1));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      if (i == j) m.at(i, j) = 0;
      else if (u(rng) < p) m.at(i, j) = w(rng);
    }
  return m;
}

static bool check_batch(const std::vector<int> &sizes, uint32_t seed) {
  const int count = (int)sizes.size();
  std::vector<Mat> mats;
  for (int j = 0; j < count; j++) mats.push_back(random_matrix(sizes[j], seed + 97 * j));
  std::vector<Mat> want;
  for (const Mat &m : mats) want.push_back(plain_fw(m));
  // the builder only stores the pointers; job j is identified by its index
  std::vector<int32_t *> ptrs(count, nullptr);
  std::vector<int32_t> n(sizes.begin(), sizes.end());
  std::vector<int64_t> ld(sizes.begin(), sizes.end());
  fapi::FwBatchDesc desc;
  fapi::build_fw_batch(count, ptrs.data(), n.data(), ld.data(), &desc);
  for (size_t L = 0; L < desc.launch_tasks.size(); L++) {
    const std::vector<Mat> snap = mats;
    std::set<std::pair<int, int64_t>> written, read_other;
    int instance = -1, tasks = 0;
    for (int q = desc.seg_off[L]; q < desc.seg_off[L + 1]; q++) {
      const fapk::BatchSeg &sg = desc.segs[q];
      const int kind = sg.kind_r & 7, r = sg.kind_r >> 3, job = sg.job;
      const int nb = desc.jobs[job].nb;
      if (instance < 0) instance = fapk::kind_instance(kind);
      if (fapk::kind_instance(kind) != instance) {
        printf("launch %zu mixes kernel instances\n", L);
        return false;
      }
      if (sg.first != tasks || sg.count != fapk::batch_task_count(kind, nb)) {
        printf("launch %zu: segment %d has first %d count %d\n", L, q, sg.first, sg.count);
        return false;
      }
      tasks += sg.count;
      const Mat &S = snap[job];
      Mat &D = mats[job];
      auto key = [&](int I, int J) { return std::make_pair(job, (int64_t)I * nb + J); };
      for (int idx = 0; idx < sg.count; idx++) {
        int I = r, J = r, p0 = r, p1 = r;
        if (kind == fapk::TASK_DIAG) {
          close_diag(D, r);
        } else if (kind == fapk::TASK_PANEL) {
          fapk::panel_tile(r, nb, idx, &I, &J);
          tile_update(S, D, I, J, r, r);
        } else if (kind == fapk::TASK_REST2) {
          fapk::pair_tile(r, nb, idx, &I, &J);
          p1 = r + 1;
          tile_update(S, D, I, J, r, r + 1);
        } else {
          int p;
          fapk::single_tile(kind, r, nb, idx, &I, &J, &p);
          p0 = p1 = p;
          tile_update(S, D, I, J, p, p);
        }
        if (I < 0 || J < 0 || I >= nb || J >= nb || !written.insert(key(I, J)).second) {
          printf("launch %zu: tile (%d, %d) of job %d out of range or written twice\n", L, I, J,
                 job);
          return false;
        }
        for (int b = p0; b <= p1 && kind != fapk::TASK_DIAG; b++) {
          if (b != J) read_other.insert(key(I, b));
          if (b != I) read_other.insert(key(b, J));
        }
      }
    }
    if (tasks != desc.launch_tasks[L]) {
      printf("launch %zu: %d tasks, descriptor says %d\n", L, tasks, desc.launch_tasks[L]);
      return false;
    }
    for (const auto &t : read_other)
      if (written.count(t)) {
        printf("launch %zu: job %d tile %lld is read and written\n", L, t.first,
               (long long)t.second);
        return false;
      }
  }
  for (int j = 0; j < count; j++)
    if (mats[j].a != want[j].a) {
      printf("job %d (n = %d) differs from plain Floyd-Warshall\n", j, sizes[j]);
      return false;
    }
  return true;
}

// ---- persistent-launch model ----------------------------------------------

struct PTask {
  int job, kind, r, idx;
  int need;  // tasks of its job that must have completed first
};

// The single-launch queue exactly as the library builds it
// (fapi::build_fw_queue), checked against the launch sequence: segments
// tile [0, T); each job's segments are its phases in launch order; each
// one's `need` is the number of its job's tasks in earlier phases.
static bool queue_list(const fapi::FwBatchDesc &desc, int count, std::vector<PTask> *out) {
  std::vector<fapk::BatchQSeg> q;
  const int32_t T = fapi::build_fw_queue(desc, &q);
  std::vector<std::vector<fapk::BatchSeg>> phases(count);
  for (size_t L = 0; L < desc.launch_tasks.size(); L++)
    for (int s = desc.seg_off[L]; s < desc.seg_off[L + 1]; s++)
      phases[desc.segs[s].job].push_back(desc.segs[s]);
  std::vector<size_t> next(count, 0);
  std::vector<int> before(count, 0);
  int32_t first = 0;
  for (size_t qi = 0; qi < q.size(); qi++) {
    const fapk::BatchQSeg &g = q[qi];
    if (g.job < 0 || g.job >= count || next[g.job] >= phases[g.job].size()) {
      printf("queue segment %zu: job %d out of range or past its phases\n", qi,
             g.job);
      return false;
    }
    const fapk::BatchSeg &want = phases[g.job][next[g.job]];
    if (g.kind_r != want.kind_r || g.count != want.count || g.first != first ||
        g.need != before[g.job]) {
      printf("queue segment %zu: job %d kind_r %d first %d count %d need %d; "
             "expected kind_r %d first %d count %d need %d\n", qi, g.job, g.kind_r,
             g.first, g.count, g.need, want.kind_r, first, want.count, before[g.job]);
      return false;
    }
    for (int idx = 0; idx < g.count; idx++)
      out->push_back(PTask{g.job, g.kind_r & 7, g.kind_r >> 3, idx, g.need});
    next[g.job]++;
    before[g.job] += g.count;
    first += g.count;
  }
  for (int j = 0; j < count; j++)
    if (next[j] != phases[j].size()) {
      printf("queue: job %d has %zu of its %zu phases\n", j, next[j],
             phases[j].size());
      return false;
    }
  if (first != T) {
    printf("queue: %d tasks, builder says %d\n", first, T);
    return false;
  }
  return true;
}

static void run_task(std::vector<Mat> &m, const fapi::FwBatchDesc &desc, const PTask &t) {
  Mat &D = m[t.job];
  const int nb = desc.jobs[t.job].nb;
  int I, J, p;
  if (t.kind == fapk::TASK_DIAG) {
    close_diag(D, t.r);
  } else if (t.kind == fapk::TASK_PANEL) {
    fapk::panel_tile(t.r, nb, t.idx, &I, &J);
    tile_update(D, D, I, J, t.r, t.r);
  } else if (t.kind == fapk::TASK_REST2) {
    fapk::pair_tile(t.r, nb, t.idx, &I, &J);
    tile_update(D, D, I, J, t.r, t.r + 1);
  } else {
    fapk::single_tile(t.kind, t.r, nb, t.idx, &I, &J, &p);
    tile_update(D, D, I, J, p, p);
  }
}

// Static: workgroup b takes tasks b, b + G, ... (the removed design).
// Claim: a running workgroup claims its next task from a shared counter
// when it has finished the current one.  ClaimAhead: as fw_queue_kernel,
// it claims the next one as soon as the current task's wait is over.
enum class Deal { Static, Claim, ClaimAhead };
enum class Outcome { Drained, Stuck, GaveUp };

// G workgroups, at most R resident; spin_limit < 0: unbounded waits.
static Outcome simulate(const std::vector<PTask> &tasks, int count,
                        const fapi::FwBatchDesc &desc, std::vector<Mat> &m, int G, int R,
                        Deal deal, long spin_limit) {
  const int T = (int)tasks.size();
  std::vector<int> done(count, 0);
  auto ready = [&](const PTask &t) { return done[t.job] >= t.need; };
  struct WG { int b = 0, task = -1, ahead = -1; long spins = 0; };
  std::vector<WG> resident;
  int next_wg = 0, claim = 0, finished = 0;
  bool flag = false;
  auto take = [&](int b, int cur) {
    if (deal == Deal::Static) {
      const long t = cur < 0 ? b : (long)cur + G;
      return t < T ? (int)t : T;
    }
    return claim < T ? claim++ : T;
  };
  while (finished < T || !resident.empty()) {
    while ((int)resident.size() < R && next_wg < G) {  // in-order dispatch into free slots
      WG w;
      w.b = next_wg++;
      w.task = take(w.b, -1);
      resident.push_back(w);
    }
    bool progress = false;
    for (size_t i = 0; i < resident.size();) {
      WG &w = resident[i];
      if (w.task < T && ready(tasks[w.task])) {
        if (deal == Deal::ClaimAhead) w.ahead = take(w.b, w.task);  // claimed before it runs
        run_task(m, desc, tasks[w.task]);
        done[tasks[w.task].job]++;
        finished++;
        w.task = deal == Deal::ClaimAhead ? w.ahead : take(w.b, w.task);
        w.spins = 0;
        progress = true;
      } else if (w.task < T) {
        if (spin_limit >= 0) {
          progress = true;  // a bounded wait runs toward its limit
          if (++w.spins > spin_limit || flag) {
            flag = true;  // gives up (or sees the flag): raises it, exits
            w.task = T;
          }
        }
      }
      if (w.task >= T) {  // exits, freeing its slot
        resident.erase(resident.begin() + i);
        progress = true;
        continue;
      }
      i++;
    }
    if (!progress) return Outcome::Stuck;
    if (resident.empty() && next_wg >= G) break;
  }
  if (flag) return Outcome::GaveUp;
  return finished == T ? Outcome::Drained : Outcome::GaveUp;
}

static bool check_persistent(const std::vector<int> &sizes, uint32_t seed) {
  const int count = (int)sizes.size();
  std::vector<Mat> mats;
  for (int j = 0; j < count; j++) mats.push_back(random_matrix(sizes[j], seed + 97 * j));
  std::vector<Mat> want;
  for (const Mat &m : mats) want.push_back(plain_fw(m));
  std::vector<int32_t *> ptrs(count, nullptr);
  std::vector<int32_t> n(sizes.begin(), sizes.end());
  std::vector<int64_t> ld(sizes.begin(), sizes.end());
  fapi::FwBatchDesc desc;
  fapi::build_fw_batch(count, ptrs.data(), n.data(), ld.data(), &desc);
  std::vector<PTask> tasks;
  if (!queue_list(desc, count, &tasks)) return false;
  const int T = (int)tasks.size();
  auto expect = [&](const char *what, int G, int R, Deal d, long lim, Outcome want_o) {
    std::vector<Mat> m = mats;
    const Outcome o = simulate(tasks, count, desc, m, G, R, d, lim);
    if (o != want_o) {
      printf("persistent model, %s (G = %d, R = %d, %d tasks): outcome %d, expected %d\n", what,
             G, R, T, (int)o, (int)want_o);
      return false;
    }
    if (o == Outcome::Drained)
      for (int j = 0; j < count; j++)
        if (m[j].a != want[j].a) {
          printf("persistent model, %s: job %d differs from plain Floyd-Warshall\n", what, j);
          return false;
        }
    return true;
  };
  const int R = 4;
  // the removed design: static deal, more workgroups than slots, unbounded waits
  if (!expect("static deal, G > R", 4 * R, R, Deal::Static, -1, Outcome::Stuck)) return false;
  // the same with a bounded wait: every workgroup exits, the flag is raised
  if (!expect("static deal, G > R, bounded wait", 4 * R, R, Deal::Static, 1000, Outcome::GaveUp))
    return false;
  // every workgroup resident
  if (!expect("static deal, G <= R", R, R, Deal::Static, -1, Outcome::Drained)) return false;
  // in-order claims, also one claim ahead (fw_queue_kernel): drain whatever the grid
  for (int G : {1, R, 4 * R, 64 * R}) {
    if (!expect("claimed in order", G, R, Deal::Claim, -1, Outcome::Drained)) return false;
    if (!expect("claimed in order, one ahead", G, R, Deal::ClaimAhead, -1, Outcome::Drained))
      return false;
    if (!expect("claimed one ahead, bounded wait", G, R, Deal::ClaimAhead, 1000000,
                Outcome::Drained))
      return false;
  }
  return true;
}

int main() {
  const std::vector<std::vector<int>> batches = {
      {1, 5, 63, 64, 65, 127, 128, 129, 200, 256, 300, 385, 449, 577},
      {640, 129, 64, 2},
      {320, 320, 321, 191},
      {0, 70},
  };
  uint32_t seed = 11;
  for (const auto &b : batches) {
    if (!check_batch(b, seed)) return 1;
    seed += 1000;
  }
  // persistent-launch model (batches whose jobs have several phases)
  const std::vector<std::vector<int>> pbatches = {{200, 129, 64}, {385, 70, 5, 128}};
  for (const auto &b : pbatches) {
    if (!check_persistent(b, seed)) return 1;
    seed += 1000;
  }
  printf("ok: %zu batches, %zu persistent models\n", batches.size(), pbatches.size());
  return 0;
}
