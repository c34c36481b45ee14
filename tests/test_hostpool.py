"""The host thread pool that spreads a scenario batch's per-scenario dense work
(`twoscalerich_amd/csrc/hostpool.h`), exercised on the CPU by a small C++ driver built
with g++: every index of a batch runs exactly once, batches follow one another without
loss, callers on several threads are served one batch at a time, and the serial limit
runs every index on the calling thread.  Also: the library's switch returns the thread
counts it promises."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "twoscalerich_amd", "csrc")

DRIVER = r"""
#include "hostpool.h"
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>
using tsr::HostPool;

static int check(int n, int reps) {
    std::vector<std::atomic<int>> hits(n > 0 ? n : 1);
    for (int r = 0; r < reps; ++r) {
        for (auto& h : hits) h.store(0);
        HostPool::get().run(n, [&](int i) { hits[i].fetch_add(1); });
        for (int i = 0; i < n; ++i)
            if (hits[i].load() != 1) { std::printf("n=%d rep=%d index %d ran %d times\n", n, r, i, hits[i].load()); return 1; }
    }
    return 0;
}

int main() {
    HostPool& p = HostPool::get();
    if (p.threads() != 4) { std::printf("threads %d\n", p.threads()); return 1; }
    for (int n : {0, 1, 2, 3, 4, 5, 17, 64, 1000})
        if (check(n, 50)) return 1;
    // two callers at once: each batch still complete
    std::atomic<int> bad{0};
    auto caller = [&](int n) { for (int r = 0; r < 200; ++r) bad += check(n, 1); };
    std::thread a(caller, 64), b(caller, 33);
    a.join(); b.join();
    if (bad.load()) return 1;
    // serial limit: every index on the calling thread, in order
    p.set_limit(1);
    if (p.threads() != 1) return 1;
    const std::thread::id me = std::this_thread::get_id();
    std::vector<int> order;
    bool elsewhere = false;
    p.run(64, [&](int i) { order.push_back(i); elsewhere |= std::this_thread::get_id() != me; });
    for (int i = 0; i < 64; ++i) if (order[i] != i) return 1;
    if (elsewhere) return 1;
    p.set_limit(0);
    if (p.threads() != 4 || check(64, 20)) return 1;
    std::printf("ok\n");
    return 0;
}
"""


def test_host_pool_runs_every_index_once(tmp_path):
    src = tmp_path / "pool.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I" + CSRC, str(src), "-o", str(exe)], check=True)
    env = dict(os.environ, TSR_HOST_THREADS="4")
    r = subprocess.run([str(exe)], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_host_threads_switch():
    from twoscalerich_amd import lib
    L = lib()
    full = L.tsr_host_threads(-1)
    assert 1 <= full <= 64
    assert L.tsr_host_threads(1) == 1
    assert L.tsr_host_threads(-1) == 1
    assert L.tsr_host_threads(0) == full
    assert L.tsr_host_threads(8) == full
