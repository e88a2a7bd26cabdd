// CPU check of the batched closure's schedule: the library's plan builder
// (fapi::build_fw_batch) and the task -> tile maps the kernels use
// (batch_desc.h) drive a host simulation of every launch on small random
// matrices.  Checked per launch: the tasks write distinct tiles, no task
// reads a tile another task of the launch writes, and every segment runs on
// the launch's one kernel instance; at the end every matrix equals plain
// Floyd-Warshall (fap/kernel/floyd.h:38-49, k-i-j order) bit for bit.
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
  printf("ok: %zu batches\n", batches.size());
  return 0;
}
