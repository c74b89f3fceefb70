// ring_model.cpp - CPU model of the device task ring (DQueue in
// psm_benchmark_amd/csrc/internal.h, claims and donation in device.hip),
// run by tests/test_ring_model.py. Test infrastructure only.
//
// Workers play the DFS lanes: they claim ring positions from `head`, wait for
// the slot's sequence word to read p + 1, read the record, release the slot
// (p + slots), and "search" the task: a task of value v is one node with two
// children of value v - 1. Children are donated to the ring while a slot is
// free (the sequence word of position tail reads tail) and searched in place
// otherwise. `busy` rises before a claim and falls when the task ends; a
// worker leaves when head == tail and busy == 0. Positions start just below
// 2^32 so that they wrap during the run. Checked: every record is read by
// exactly the claimant of its position (a tag equal to the position), and
// the nodes searched equal the closed form sum over the initial tasks of
// 2^(v+1) - 1.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

struct Rec {
    std::atomic<uint32_t> value{0};
    std::atomic<uint32_t> tag{0};
};

int main(int argc, char** argv) {
    const uint32_t slots = argc > 1 ? (uint32_t)atoi(argv[1]) : 8;   // power of two
    const uint32_t workers = argc > 2 ? (uint32_t)atoi(argv[2]) : 8;
    const uint32_t depth = argc > 3 ? (uint32_t)atoi(argv[3]) : 14;
    const uint32_t start = 0xFFFFFFFFu - 2u * slots;   // wraps early
    const uint32_t mask = slots - 1;
    std::vector<Rec> ring(slots);
    std::vector<std::atomic<uint32_t>> seq(slots);
    std::atomic<uint32_t> head{start}, tail{start}, busy{0};
    std::atomic<uint64_t> nodes{0}, bad{0}, donated{0}, skipped{0};
    const uint32_t n_init = slots / 2;
    // seq words are indexed by slot: position p lives in slot p & mask
    for (uint32_t i = 0; i < slots; ++i) {
        const uint32_t p = start + i;
        seq[p & mask].store(p, std::memory_order_relaxed);
    }
    for (uint32_t k = 0; k < n_init; ++k) {
        const uint32_t p = start + k;
        ring[p & mask].value.store(depth, std::memory_order_relaxed);
        ring[p & mask].tag.store(p, std::memory_order_relaxed);
        seq[p & mask].store(p + 1, std::memory_order_release);
    }
    tail.store(start + n_init);

    auto worker = [&]() {
        std::vector<uint32_t> stack;
        while (true) {
            uint32_t h = head.load(), t = tail.load();
            if (t == h) {
                if (head.load() == tail.load() && busy.load() == 0) return;
                std::this_thread::yield();
                continue;
            }
            busy.fetch_add(1);
            if (!head.compare_exchange_strong(h, h + 1)) {
                busy.fetch_sub(1);
                continue;
            }
            while (seq[h & mask].load(std::memory_order_acquire) != h + 1) std::this_thread::yield();
            const uint32_t v = ring[h & mask].value.load(std::memory_order_relaxed);
            if (ring[h & mask].tag.load(std::memory_order_relaxed) != h) bad.fetch_add(1);
            seq[h & mask].store(h + slots, std::memory_order_release);
            stack.assign(1, v);
            while (!stack.empty()) {
                const uint32_t x = stack.back();
                stack.pop_back();
                nodes.fetch_add(1, std::memory_order_relaxed);
                if (x == 0) continue;
                for (int c = 0; c < 2; ++c) {
                    uint32_t tt = tail.load();
                    if (seq[tt & mask].load(std::memory_order_acquire) == tt &&
                        tail.compare_exchange_strong(tt, tt + 1)) {
                        ring[tt & mask].value.store(x - 1, std::memory_order_relaxed);
                        ring[tt & mask].tag.store(tt, std::memory_order_relaxed);
                        seq[tt & mask].store(tt + 1, std::memory_order_release);
                        donated.fetch_add(1);
                    } else {
                        skipped.fetch_add(1);
                        stack.push_back(x - 1);
                    }
                }
            }
            busy.fetch_sub(1);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < workers; ++w) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    const uint64_t expect = (uint64_t)n_init * ((2ull << depth) - 1);
    printf("{\"nodes\": %llu, \"expect\": %llu, \"bad\": %llu, \"donated\": %llu, \"skipped\": %llu, "
           "\"head\": %u, \"tail\": %u, \"wrapped\": %d}\n",
           (unsigned long long)nodes.load(), (unsigned long long)expect, (unsigned long long)bad.load(),
           (unsigned long long)donated.load(), (unsigned long long)skipped.load(), head.load(), tail.load(),
           head.load() < start ? 1 : 0);
    return nodes.load() == expect && bad.load() == 0 && head.load() == tail.load() ? 0 : 1;
}
