// Work-item scheduler of the host kernel emulation (test infrastructure
// only): each workgroup's work-items run as fibers (ucontext) on one host
// thread, switching at barriers.  With EMU_THREADS=n (n > 1) a launch runs on
// n host threads that take workgroups in increasing order from a shared
// counter, as the hardware dispatcher does, so workgroups overlap in time and
// cross-workgroup protocols (tickets, look-back status words) meet real
// concurrency; `__shared__` storage and the scheduler state are per thread.
#include <hip/hip_runtime.h>

#include <ucontext.h>

#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

namespace emu {
namespace {

constexpr size_t kStack = 256 * 1024;

thread_local BlockSync g_sync;

struct Fiber {
  ucontext_t uc;
  std::unique_ptr<char[]> stack;
  Ctx ctx;
  Bar* wait_bar = nullptr;
  uint64_t wait_gen = 0;
  bool done = false;
};

struct Sched {
  ucontext_t main;
  std::vector<Fiber> fibers;
  int cur = -1;
  const std::function<void()>* fn = nullptr;
  void* main_fake_stack = nullptr;
  const void* main_stack_bottom = nullptr;
  size_t main_stack_size = 0;
};
thread_local Sched S;

void finish_switch_into_fiber() {
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &S.main_stack_bottom, &S.main_stack_size);
#endif
}

// Switches from the running fiber back to the scheduler.
void yield_to_main(bool exiting) {
  Fiber& f = S.fibers[S.cur];
  f.ctx = ctx;
#if EMU_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(exiting ? nullptr : &fake, S.main_stack_bottom,
                                 S.main_stack_size);
  swapcontext(&f.uc, &S.main);
  __sanitizer_finish_switch_fiber(fake, &S.main_stack_bottom, &S.main_stack_size);
#else
  (void)exiting;
  swapcontext(&f.uc, &S.main);
#endif
  ctx = f.ctx;
}

void fiber_entry() {
  finish_switch_into_fiber();
  Fiber& f = S.fibers[S.cur];
  ctx = f.ctx;
  (*S.fn)();
  // The work-item has returned: it no longer takes part in barriers.
  g_sync.alive[ctx.wave][ctx.lane] = false;
  for (Bar* b : {&g_sync.wave[ctx.wave], &g_sync.block}) {
    b->expected -= 1;
    if (b->arrived > 0 && b->arrived == b->expected) {
      b->arrived = 0;
      b->gen += 1;
    }
  }
  S.fibers[S.cur].done = true;
  yield_to_main(true);
}

void switch_to(int i) {
  S.cur = i;
  Fiber& f = S.fibers[i];
  ctx = f.ctx;
#if EMU_ASAN
  __sanitizer_start_switch_fiber(&S.main_fake_stack, f.stack.get(), kStack);
  swapcontext(&S.main, &f.uc);
  __sanitizer_finish_switch_fiber(S.main_fake_stack, nullptr, nullptr);
#else
  swapcontext(&S.main, &f.uc);
#endif
}

// Runs workgroup (bx, by) to completion on the calling thread.
void run_block(dim3 grid, dim3 block, unsigned bx, unsigned by) {
  const unsigned n = block.x * block.y * block.z;
  const unsigned nwaves = n / 64;
  if (S.fibers.size() < n) S.fibers.resize(n);
  for (unsigned i = 0; i < n; ++i)
    if (!S.fibers[i].stack) S.fibers[i].stack.reset(new char[kStack]);
  g_sync.block = Bar{(int)n, 0, 0};
  for (unsigned w = 0; w < nwaves; ++w) {
    g_sync.wave[w] = Bar{64, 0, 0};
    for (int l = 0; l < 64; ++l) g_sync.alive[w][l] = true;
  }
  for (unsigned i = 0; i < n; ++i) {
    Fiber& f = S.fibers[i];
    f.ctx.tid = {i % block.x, (i / block.x) % block.y, 0};
    f.ctx.bid = {bx, by, 0};
    f.ctx.bdim = block;
    f.ctx.gdim = grid;
    f.ctx.lane = (int)(i & 63);
    f.ctx.wave = (int)(i >> 6);
    f.wait_bar = nullptr;
    f.done = false;
    getcontext(&f.uc);
    f.uc.uc_stack.ss_sp = f.stack.get();
    f.uc.uc_stack.ss_size = kStack;
    f.uc.uc_link = nullptr;
    makecontext(&f.uc, fiber_entry, 0);
  }
  unsigned live = n;
  while (live > 0) {
    bool progress = false;
    for (unsigned i = 0; i < n; ++i) {
      Fiber& f = S.fibers[i];
      if (f.done) continue;
      if (f.wait_bar && f.wait_bar->gen == f.wait_gen) continue;
      f.wait_bar = nullptr;
      switch_to((int)i);
      progress = true;
      if (f.done) --live;
    }
    if (!progress) {
      fprintf(stderr, "emu: deadlock in block (%u, %u): no work-item can proceed\n", bx, by);
      abort();
    }
  }
}

int launch_threads() {
  const char* e = getenv("EMU_THREADS");
  const int t = e ? atoi(e) : 1;
  return t < 1 ? 1 : (t > 64 ? 64 : t);
}

}  // namespace

BlockSync* sync_state() { return &g_sync; }

void bar_wait(Bar* b) {
  b->arrived += 1;
  if (b->arrived == b->expected) {
    b->arrived = 0;
    b->gen += 1;
    return;
  }
  Fiber& f = S.fibers[S.cur];
  f.wait_bar = b;
  f.wait_gen = b->gen;
  yield_to_main(false);
}

void pause() { std::this_thread::yield(); }

void run_grid(dim3 grid, dim3 block, const std::function<void()>& fn) {
  const unsigned n = block.x * block.y * block.z;
  if (n == 0 || n > (unsigned)kMaxThreads || (n % 64) != 0) {
    fprintf(stderr, "emu: unsupported block size %u\n", n);
    abort();
  }
  const unsigned long long nblocks = (unsigned long long)grid.x * grid.y;
  const int threads = launch_threads();
  if (threads == 1 || nblocks <= 1) {
    const std::function<void()>* saved_fn = S.fn;
    S.fn = &fn;
    for (unsigned by = 0; by < grid.y; ++by)
      for (unsigned bx = 0; bx < grid.x; ++bx) run_block(grid, block, bx, by);
    S.fn = saved_fn;
    return;
  }
  std::atomic<unsigned long long> next{0};
  auto worker = [&]() {
    S.fn = &fn;
    for (;;) {
      const unsigned long long b = next.fetch_add(1);
      if (b >= nblocks) break;
      run_block(grid, block, (unsigned)(b % grid.x), (unsigned)(b / grid.x));
    }
  };
  std::vector<std::thread> pool;
  const int t = (unsigned long long)threads < nblocks ? threads : (int)nblocks;
  for (int i = 0; i < t; ++i) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
}

}  // namespace emu
