// The library's context-free host entry points (scene generators, kd-tree
// build, query and export, camera rays, Perlin table, batch item plan) called
// from several threads at once, for a build under ThreadSanitizer.  Each
// thread's results must equal the same work done on one thread; the
// sanitizer aborts on a data race between the calls.  Built without OpenMP:
// the race of interest is between independent calls (the Python tuners build
// trees from a thread pool), and an OpenMP runtime built without the
// sanitizer would report its own internal synchronisation.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "rayengine.h"

namespace {

struct Result {
  std::vector<unsigned char> nodes, leaves, tris;
  std::vector<float> o, d;
  uint32_t perm[512];
  std::vector<unsigned char> items;
  int64_t n_items = -1;
  bool ok = true;
};

Result work(int k) {
  Result r;
  const int64_t n = 1500 + 250 * k;
  std::vector<float> soup((size_t)n * 9);
  r.ok &= re_gen_cluster_soup(100 + k, n, 8, -2.0f, 2.0f, 0.3f, -3.0f, 0.5f, soup.data()) == RE_OK;
  re_kd_build_params p;
  re_kd_default_params(&p);
  if (k % 2) {  // the reference's spatial-median rule on every other thread
    p.split_rule = 1;
    p.clip = 0;
    p.max_leaf_tris = 12;
  }
  re_kdtree* t = nullptr;
  r.ok &= re_kd_build(soup.data(), n, &p, &t) == RE_OK;
  re_kd_info info;
  r.ok &= t && re_kd_get_info(t, &info) == RE_OK;
  if (r.ok) {
    r.nodes.resize((size_t)info.n_nodes * 8 + 16);
    r.leaves.resize((size_t)info.n_leaves * 64);
    r.tris.resize((size_t)info.n_trirecs * 40);
    r.ok &= re_kd_export(t, r.nodes.data(), r.leaves.data(), r.tris.data()) == RE_OK;
  }
  re_kd_free(t);
  re_camera cam;
  std::memset(&cam, 0, sizeof(cam));
  cam.pos[2] = 5.0f + k;
  cam.view[2] = -1.0f;
  cam.camdown[1] = -1.0f;
  cam.fov[0] = cam.fov[1] = 60.0f;
  cam.focal_distance = 1.0f;
  r.o.resize(3 * 48 * 32);
  r.d.resize(3 * 48 * 32);
  r.ok &= re_gen_camera_rays(&cam, 48, 32, r.o.data(), r.d.data()) == RE_OK;
  r.ok &= re_perlin_table(2016 + k, r.perm) == RE_OK;
  re_batch_scene sc[3] = {{nullptr, 0, 1000 + k}, {nullptr, 1000 + k, 37}, {nullptr, 1037 + k, 500}};
  re_batch_tile tl[1] = {{-1, 0, 1537 + k, 300, -1, 0}};
  r.ok &= re_batch_plan(sc, 3, tl, 1, 1837 + k, 64, nullptr, 0, &r.n_items, nullptr) == RE_OK;
  r.items.resize((size_t)r.n_items * 24);
  r.ok &= re_batch_plan(sc, 3, tl, 1, 1837 + k, 64, r.items.data(), r.n_items, &r.n_items, nullptr) == RE_OK;
  // a refused call records its message in the calling thread's error slot
  r.ok &= re_kd_build(soup.data(), -1, nullptr, &t) != RE_OK;
  return r;
}

bool same(const Result& a, const Result& b) {
  return a.ok && b.ok && a.nodes == b.nodes && a.leaves == b.leaves && a.tris == b.tris && a.o == b.o &&
         a.d == b.d && std::memcmp(a.perm, b.perm, sizeof(a.perm)) == 0 && a.items == b.items &&
         a.n_items == b.n_items;
}

}  // namespace

int main() {
  constexpr int kThreads = 6;
  std::vector<Result> ref;
  for (int k = 0; k < kThreads; ++k) ref.push_back(work(k));
  int failed = 0;
  for (int round = 0; round < 2; ++round) {
    std::vector<Result> got(kThreads);
    std::vector<std::thread> th;
    for (int k = 0; k < kThreads; ++k) th.emplace_back([&got, k] { got[k] = work(k); });
    for (auto& x : th) x.join();
    for (int k = 0; k < kThreads; ++k)
      if (!same(ref[k], got[k])) {
        std::fprintf(stderr, "FAILED: thread %d of round %d differs from the one-thread result\n", k, round);
        ++failed;
      }
  }
  std::printf("%d failed checks\n", failed);
  return failed ? 1 : 0;
}
