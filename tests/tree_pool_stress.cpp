// Stress of the host tree worker pool (qz_tree.cpp): jobs of ragged sizes from
// two calling threads at once; every index of every job must run exactly once.
// Built and run by test_tree_pool.py.
#include "../quantizer_amd/csrc/qz_tree.cpp"

#include <cstdio>

int main()
{
  using namespace qz;
  std::atomic<long long> bad(0);
  auto caller = [&](int seed) {
    for (int it = 0; it < 2000; it++)
    {
      const int count = (it * 7919 + seed) % 300;
      std::vector<std::atomic<int>> hit(count);
      for (auto &h : hit)
        h.store(0);
      parallel_for(count, [&](int q) { hit[q]++; });
      for (int q = 0; q < count; q++)
        if (hit[q].load() != 1)
          bad++;
    }
  };
  std::thread a(caller, 1), b(caller, 2);
  a.join();
  b.join();
  std::printf("workers %d, wrong counts %lld\n", TreePool::get().workers(), bad.load());
  return bad.load() == 0 ? 0 : 1;
}
