// Stress driver for p3g::HostPool (host_pool.hpp), built with g++ by
// test_host_pool.py.  Exits 0 when every check holds, prints the failure
// otherwise.
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "host_pool.hpp"

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);    \
      fails++;                                                    \
    }                                                             \
  } while (0)

int main() {
  p3g::HostPool pool(8);
  CHECK(pool.threads() == 8);

  // every index exactly once, sizes 0..300
  for (size_t n = 0; n <= 300; n += 7) {
    std::vector<std::atomic<int>> hit(n);
    for (auto& h : hit) h = 0;
    pool.run(n, [&](size_t i) { hit[i]++; });
    for (size_t i = 0; i < n; i++) CHECK(hit[i] == 1);
  }

  // concurrent callers: 6 threads x 200 loops each, all results exact
  {
    std::atomic<int> bad{0};
    std::vector<std::thread> cs;
    for (int c = 0; c < 6; c++)
      cs.emplace_back([&, c] {
        for (int it = 0; it < 200; it++) {
          const size_t n = 1 + (size_t)((c * 37 + it * 11) % 97);
          std::vector<uint64_t> out(n, 0);
          pool.run(n, [&](size_t i) { out[i] = i * i + (uint64_t)c; });
          for (size_t i = 0; i < n; i++)
            if (out[i] != i * i + (uint64_t)c) bad++;
        }
      });
    for (auto& t : cs) t.join();
    CHECK(bad == 0);
  }

  // nested loops from inside a loop body
  {
    std::atomic<uint64_t> sum{0};
    pool.run(16, [&](size_t i) { pool.run(32, [&](size_t j) { sum += i * 32 + j; }); });
    CHECK(sum == (uint64_t)(512 * 511 / 2));
  }

  // an exception reaches the caller after every other index has run
  {
    std::atomic<int> ran{0};
    bool thrown = false;
    try {
      pool.run(100, [&](size_t i) {
        ran++;
        if (i == 37) throw std::runtime_error("boom");
      });
    } catch (const std::runtime_error&) {
      thrown = true;
    }
    CHECK(thrown);
    CHECK(ran == 100);
    // the pool still works afterwards
    std::atomic<int> cnt{0};
    pool.run(50, [&](size_t) { cnt++; });
    CHECK(cnt == 50);
  }

  // a one-thread pool runs loops inline
  {
    p3g::HostPool one(1);
    std::vector<int> v(10, 0);
    one.run(10, [&](size_t i) { v[i] = (int)i; });
    for (int i = 0; i < 10; i++) CHECK(v[i] == i);
  }

  if (fails) return 1;
  std::printf("host pool ok\n");
  return 0;
}
