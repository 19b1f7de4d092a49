// gtap_protocol_model.cpp -- CPU model of the runtime's lock-free protocols, run under
// ThreadSanitizer by tests/test_protocol_model.py. Each model restates one protocol of the
// device code with std::atomic and checks its invariant over many randomized interleavings:
//
//  deque   the (top << 32 | bottom) deque word of gtap_thread.hpp (try_pop, try_steal, the
//          epilogue's publish): one owner appends, keeps up to 64 newest, publishes the rest
//          with one add, pops up to 64 newest with one CAS; thieves take half (<= 64) of the
//          oldest with one CAS and read the slots after it. The owner never writes within
//          kMargin of the top it last saw. Invariant: every task id is taken exactly once.
//  ring    the id rings of 64-id cells with Vyukov sequence numbers (ring_put_claim,
//          ring_probe / refill_chunk): producers take an enqueue ticket and wait for the cell,
//          consumers CAS the dequeue position of a full cell. Invariant: every id once.
//  quiesce the all-tasks termination rule (thread_scheduler's idle path): idle count compared
//          with the started count, job owners counted as started by the reset. Invariant:
//          done is never set while a task is outstanding, also when the job owner starts
//          after every other worker went idle; the former rule (owners count themselves when
//          they start) is shown to end early in that schedule. Also run with the idle count
//          joined only after 4 consecutive failed acquire passes (kIdleRegisterAfter).
//  ovf     the deque overflow rings (gtap_thread.hpp ovf_fill / ovf_pop / ovf_spill_deque,
//          ovf_drain_kernel in gtap_host.hip): cells of state | position tag | id, a ticket
//          fetch-add per batch, producers wait for the cell of the previous lap to be freed,
//          consumers CAS the head over a run of READY cells and free them for the next lap;
//          the drain frees what an aborted launch left between head and tickets. Invariant:
//          every id once, over many laps of a small ring, and a ring that works after a drain.
//  ovfq    quiescence with the rings (thread_scheduler's idle path): a producer bumps every
//          group's epoch after a push; a worker looks at the ring only when its group's epoch
//          moved since it last found the ring empty, and counts itself idle only when it has
//          not moved. Invariant: done is never set while a task is outstanding (in a worker or
//          the ring); without the epoch rule, a worker that counts itself idle before its next
//          look ends the launch early.
//  join    the additive join word (gtap_dev.hpp: count [0..11] | resume queue [12..15] |
//          signed sum [16..63]): children add (value << 16) - 1 concurrently. Invariant: the
//          last child alone sees count 1, the queue bits survive, the final sum is exact at
//          the bounds (4095 children, |value| < 2^35).
//
// The GPU publishes with "complete, then atomic" (s_waitcnt vmcnt(0) before the publishing
// atomic) and reads after the atomic that handed it the work returned; the model writes those
// as release on the publishing atomic and acquire on the claiming one.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

static int g_failures = 0;
#define CHECK(c, ...)                                 \
    do {                                              \
        if (!(c)) {                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);             \
            fprintf(stderr, "\n");                    \
            ++g_failures;                             \
        }                                             \
    } while (0)

// ------------------------------------------------------------------ deque word
namespace deque_model {

constexpr uint32_t kCap = 4096, kMargin = 1024, kWave = 64;

struct Deque {
    std::atomic<uint64_t> word{0};
    std::atomic<uint32_t> slots[kCap];
    Deque() {
        for (auto& s : slots) s.store(0, std::memory_order_relaxed);
    }
};

void run(unsigned seed, int thieves, uint32_t total) {
    Deque dq;
    std::vector<std::atomic<uint8_t>> taken(total + 1);
    for (auto& t : taken) t.store(0);
    std::atomic<bool> owner_done{false};
    std::atomic<uint32_t> ntaken{0};
    auto take = [&](uint32_t id) {
        CHECK(id >= 1 && id <= total, "deque: id %u out of range", id);
        if (id >= 1 && id <= total) {
            const uint8_t before = taken[id].fetch_add(1);
            CHECK(before == 0, "deque: id %u taken twice", id);
        }
        ntaken.fetch_add(1);
    };
    std::thread owner([&] {
        std::mt19937 rng(seed);
        uint32_t next_id = 1, tail = 0, bottom = 0, top_cache = 0;
        while (next_id <= total) {
            // a round appends some children privately (the owner's end of the ring)
            uint32_t n = rng() % 130;
            if (n > total + 1 - next_id) n = total + 1 - next_id;
            if (tail + n - top_cache > kCap - kMargin) {
                top_cache = (uint32_t)(dq.word.load(std::memory_order_relaxed) >> 32);
                if (tail + n - top_cache > kCap - kMargin) {
                    n = 0;  // full: the round keeps what it has (the device reports overflow)
                }
            }
            for (uint32_t i = 0; i < n; ++i) dq.slots[(tail + i) & (kCap - 1)].store(next_id + i, std::memory_order_relaxed);
            tail += n;
            next_id += n;
            // keep up to 64 newest for the next round, publish the rest with one add
            uint32_t keep = tail - bottom < kWave ? tail - bottom : kWave;
            keep = rng() % (keep + 1);
            for (uint32_t i = tail - keep; i < tail; ++i) take(dq.slots[i & (kCap - 1)].load(std::memory_order_relaxed));
            tail -= keep;
            if (tail != bottom) {
                const uint64_t old = dq.word.fetch_add(tail - bottom, std::memory_order_release);
                top_cache = (uint32_t)(old >> 32);
                bottom = tail;
            }
            // sometimes pop up to 64 of the newest published (the next round's batch)
            if (rng() % 3 == 0) {
                while (true) {
                    uint64_t wv = dq.word.load(std::memory_order_relaxed);
                    const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
                    const int32_t sz = (int32_t)(b - t);
                    if (sz <= 0) break;
                    const uint32_t k = (uint32_t)sz < kWave ? (uint32_t)sz : kWave;
                    if (dq.word.compare_exchange_strong(wv, ((uint64_t)t << 32) | (b - k), std::memory_order_acquire,
                                                        std::memory_order_relaxed)) {
                        for (uint32_t i = b - k; i < b; ++i) take(dq.slots[i & (kCap - 1)].load(std::memory_order_relaxed));
                        bottom = tail = b - k;
                        top_cache = t;
                        break;
                    }
                }
            }
        }
        // drain: pop until empty (a failed CAS means a thief took tasks; retry)
        while (true) {
            uint64_t wv = dq.word.load(std::memory_order_relaxed);
            const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
            const int32_t sz = (int32_t)(b - t);
            if (sz <= 0) break;
            const uint32_t k = (uint32_t)sz < kWave ? (uint32_t)sz : kWave;
            if (dq.word.compare_exchange_strong(wv, ((uint64_t)t << 32) | (b - k), std::memory_order_acquire,
                                                std::memory_order_relaxed))
                for (uint32_t i = b - k; i < b; ++i) take(dq.slots[i & (kCap - 1)].load(std::memory_order_relaxed));
        }
        owner_done.store(true);
    });
    std::vector<std::thread> th;
    for (int v = 0; v < thieves; ++v) {
        th.emplace_back([&, v] {
            std::mt19937 rng(seed * 31 + v);
            while (!(owner_done.load() && ntaken.load() == total)) {
                uint64_t wv = dq.word.load(std::memory_order_relaxed);
                const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
                const int32_t sz = (int32_t)(b - t);
                if (sz <= 0) {
                    if (owner_done.load()) break;
                    std::this_thread::yield();
                    continue;
                }
                uint32_t k = ((uint32_t)sz + 1) / 2;
                const uint32_t batch = 1 + rng() % kWave;
                if (k > batch) k = batch;
                if (dq.word.compare_exchange_strong(wv, wv + ((uint64_t)k << 32), std::memory_order_acquire,
                                                    std::memory_order_relaxed))
                    for (uint32_t i = t; i < t + k; ++i) take(dq.slots[i & (kCap - 1)].load(std::memory_order_relaxed));
            }
        });
    }
    owner.join();
    for (auto& t : th) t.join();
    uint32_t missing = 0;
    for (uint32_t i = 1; i <= total; ++i) missing += taken[i].load() == 0;
    CHECK(missing == 0 && ntaken.load() == total, "deque: %u ids never taken, %u takes for %u ids", missing,
          ntaken.load(), total);
}

}  // namespace deque_model

// ------------------------------------------------------------------ id ring (Vyukov cells)
namespace ring_model {

constexpr uint32_t kChunk = 64;

struct Ring {
    uint32_t ncells;
    std::atomic<uint32_t> enq{0}, deq{0};
    std::vector<std::atomic<uint32_t>> seq;
    std::vector<std::atomic<uint32_t>> ids;
    explicit Ring(uint32_t n) : ncells(n), seq(n), ids((size_t)n * kChunk) {
        for (uint32_t i = 0; i < n; ++i) seq[i].store(i);
    }
    void put(const uint32_t* v) {
        const uint32_t pos = enq.fetch_add(1, std::memory_order_relaxed);  // ticket
        const uint32_t c = pos & (ncells - 1);
        while (seq[c].load(std::memory_order_acquire) != pos) std::this_thread::yield();  // the cell is free
        for (uint32_t i = 0; i < kChunk; ++i) ids[(size_t)c * kChunk + i].store(v[i], std::memory_order_relaxed);
        seq[c].store(pos + 1, std::memory_order_release);  // full
    }
    bool take(uint32_t* out) {
        for (int tries = 0; tries < 4; ++tries) {
            uint32_t pos = deq.load(std::memory_order_relaxed);
            const uint32_t c = pos & (ncells - 1);
            const int32_t diff = (int32_t)(seq[c].load(std::memory_order_acquire) - (pos + 1));
            if (diff == 0) {
                if (deq.compare_exchange_strong(pos, pos + 1, std::memory_order_acquire, std::memory_order_relaxed)) {
                    for (uint32_t i = 0; i < kChunk; ++i) out[i] = ids[(size_t)c * kChunk + i].load(std::memory_order_relaxed);
                    seq[c].store(pos + ncells, std::memory_order_release);  // free for the next lap
                    return true;
                }
            } else if (diff < 0) {
                return false;  // empty
            }
        }
        return false;
    }
};

void run(unsigned seed, int producers, int consumers, uint32_t cells_each, uint32_t ncells) {
    Ring ring(ncells);
    const uint32_t total = (uint32_t)producers * cells_each * kChunk;
    std::vector<std::atomic<uint8_t>> seen(total);
    for (auto& s : seen) s.store(0);
    std::atomic<uint32_t> got{0};
    std::vector<std::thread> th;
    for (int p = 0; p < producers; ++p)
        th.emplace_back([&, p] {
            uint32_t v[kChunk];
            for (uint32_t c = 0; c < cells_each; ++c) {
                for (uint32_t i = 0; i < kChunk; ++i) v[i] = ((uint32_t)p * cells_each + c) * kChunk + i;
                ring.put(v);
            }
        });
    for (int q = 0; q < consumers; ++q)
        th.emplace_back([&, q] {
            std::mt19937 rng(seed + q);
            uint32_t v[kChunk];
            while (got.load() < total) {
                if (ring.take(v)) {
                    for (uint32_t i = 0; i < kChunk; ++i) {
                        CHECK(v[i] < total, "ring: id %u out of range", v[i]);
                        if (v[i] < total) CHECK(seen[v[i]].fetch_add(1) == 0, "ring: id %u twice", v[i]);
                    }
                    got.fetch_add(kChunk);
                } else if (rng() % 4 == 0) {
                    std::this_thread::yield();
                }
            }
        });
    for (auto& t : th) t.join();
    uint32_t missing = 0;
    for (uint32_t i = 0; i < total; ++i) missing += seen[i].load() == 0;
    CHECK(missing == 0, "ring: %u ids lost", missing);
}

}  // namespace ring_model

// ------------------------------------------------------------------ deque overflow ring
namespace ovf_model {

constexpr uint64_t kReady = 2ull << 62;
constexpr uint64_t kTag = 0x3FFFFFFFull;
uint64_t free_word(uint64_t pos) { return (pos & kTag) << 32; }
uint64_t ready_word(uint64_t pos, uint32_t id) { return kReady | ((pos & kTag) << 32) | id; }

struct Ring {
    const uint32_t cap;
    std::vector<std::atomic<uint64_t>> cells;
    std::atomic<uint64_t> head{0}, tickets{0};
    explicit Ring(uint32_t c) : cap(c), cells(c) {
        for (uint32_t i = 0; i < c; ++i) cells[i].store(free_word(i));
    }
    std::atomic<uint64_t>& cell(uint64_t pos) { return cells[pos & (cap - 1)]; }
    // ovf_fill: wait until the previous lap's consumer freed the cell, then publish
    void fill(uint64_t pos, uint32_t id) {
        while (cell(pos).load(std::memory_order_acquire) != free_word(pos)) std::this_thread::yield();
        cell(pos).store(ready_word(pos, id), std::memory_order_release);
    }
    // the producers' room check (ovf_room; kMargin = 0 here)
    bool room(uint64_t base, uint32_t n) { return base + n - head.load() <= cap; }
    // ovf_pop with `lanes` lanes: the ready run at the head, one CAS
    uint32_t pop(uint32_t lanes, uint32_t* out) {
        const uint64_t h = head.load(), tk = tickets.load();
        if ((int64_t)(tk - h) <= 0) return 0;
        const uint32_t avail = (uint32_t)std::min<uint64_t>(tk - h, lanes);
        uint64_t cv[64];
        uint32_t k = 0;
        while (k < avail) {
            cv[k] = cell(h + k).load(std::memory_order_acquire);
            if (cv[k] != ready_word(h + k, (uint32_t)cv[k])) break;
            ++k;
        }
        if (k == 0) return 0;
        uint64_t expected = h;
        if (!head.compare_exchange_strong(expected, h + k, std::memory_order_acq_rel)) return 0;
        for (uint32_t i = 0; i < k; ++i) {
            out[i] = (uint32_t)cv[i];
            cell(h + i).store(free_word(h + i + cap), std::memory_order_release);
        }
        return k;
    }
    // ovf_drain_kernel, with no launch running
    void drain() {
        const uint64_t h = head.load(), tk = tickets.load();
        for (uint64_t p = h; p < tk; ++p) cell(p).store(free_word(p + cap));
        head.store(tk);
    }
};

// producers push batches of up to 16 ids (throttled to keep half the ring free, where the
// runtime sizes the ring instead), consumers pop with 1 to 8 lanes; the ring laps many times
void run(unsigned seed, int producers, int consumers, uint32_t per_producer, uint32_t cap) {
    Ring ring(cap);
    const uint32_t total = (uint32_t)producers * per_producer;
    std::vector<std::atomic<uint8_t>> seen(total);
    for (auto& x : seen) x.store(0);
    std::atomic<uint32_t> got{0};
    std::vector<std::thread> th;
    for (int p = 0; p < producers; ++p)
        th.emplace_back([&, p] {
            std::mt19937 rng(seed * 31 + p);
            uint32_t next = 0;
            while (next < per_producer) {
                const uint32_t k = std::min<uint32_t>(1 + rng() % 16, per_producer - next);
                if (ring.tickets.load() - ring.head.load() + k > cap / 2) {
                    std::this_thread::yield();
                    continue;
                }
                const uint64_t base = ring.tickets.fetch_add(k);
                CHECK(ring.room(base, k), "ovf: positions beyond the head's lap");
                for (uint32_t j = 0; j < k; ++j) ring.fill(base + j, (uint32_t)p * per_producer + next + j);
                next += k;
            }
        });
    for (int c = 0; c < consumers; ++c)
        th.emplace_back([&, c] {
            std::mt19937 rng(seed * 17 + c);
            uint32_t out[64];
            while (got.load() < total) {
                const uint32_t k = ring.pop(1 + rng() % 8, out);
                for (uint32_t i = 0; i < k; ++i) {
                    CHECK(out[i] < total, "ovf: id %u out of range", out[i]);
                    if (out[i] < total) CHECK(seen[out[i]].fetch_add(1) == 0, "ovf: id %u twice", out[i]);
                }
                got.fetch_add(k);
                if (k == 0) std::this_thread::yield();
            }
        });
    for (auto& t : th) t.join();
    uint32_t missing = 0;
    for (uint32_t i = 0; i < total; ++i) missing += seen[i].load() == 0;
    CHECK(missing == 0, "ovf: %u ids lost", missing);
    CHECK(ring.head.load() == ring.tickets.load(), "ovf: head %llu, tickets %llu", (unsigned long long)ring.head.load(),
          (unsigned long long)ring.tickets.load());

    // an aborted launch: tickets taken and some cells filled, none claimed; the drain frees
    // them, and the next launch's pushes and pops over the same cells see every id once
    const uint64_t base = ring.tickets.fetch_add(cap / 2);
    for (uint32_t j = 0; j < cap / 4; ++j) ring.fill(base + j, 7);
    ring.drain();
    CHECK(ring.head.load() == ring.tickets.load(), "ovf: drain left tickets");
    const uint32_t again = 3 * cap;
    std::vector<uint8_t> seen2(again, 0);
    uint32_t got2 = 0, out[64];
    for (uint32_t i = 0; i < again; ++i) {
        const uint64_t b = ring.tickets.fetch_add(1);
        CHECK(ring.room(b, 1), "ovf: no room after the drain");
        ring.fill(b, i);
        if (i % 3 == 2)
            for (uint32_t k; (k = ring.pop(8, out)) != 0;) {
                for (uint32_t x = 0; x < k; ++x) seen2[out[x]] += 1;
                got2 += k;
            }
    }
    for (uint32_t k; (k = ring.pop(8, out)) != 0;) {
        for (uint32_t x = 0; x < k; ++x) seen2[out[x]] += 1;
        got2 += k;
    }
    uint32_t bad = 0;
    for (uint32_t i = 0; i < again; ++i) bad += seen2[i] != 1;
    CHECK(got2 == again && bad == 0, "ovf: after the drain %u of %u ids, %u wrong", got2, again, bad);
}

}  // namespace ovf_model

// ------------------------------------------------------------------ quiescence
namespace quiesce_model {

// Workers 0 .. owners-1 start one job each by index; a job is `work` tasks that its owner
// runs and may publish to a shared pool from which idle workers take. `owner_delay` makes
// the owners start late (not yet resident); a worker joins the idle count only after
// `register_after` consecutive failed acquire passes. Returns true if done was set while tasks
// were still outstanding.
bool run(unsigned seed, int workers, int owners, int work, bool count_owners_at_reset, int owner_delay_us,
         int register_after = 0) {
    std::atomic<int> idle{0}, started{count_owners_at_reset ? owners : 0}, done{0};
    std::atomic<int> pool{0};         // published tasks
    std::atomic<int> outstanding{owners * work};
    std::atomic<bool> early{false};
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w) {
        th.emplace_back([&, w] {
            std::mt19937 rng(seed * 7 + w);
            const bool owner = w < owners;
            if (owner && owner_delay_us) std::this_thread::sleep_for(std::chrono::microseconds(owner_delay_us));
            if (!count_owners_at_reset || !owner) started.fetch_add(1);
            int mine = owner ? work : 0;  // tasks this worker holds (its job, then taken ones)
            bool registered = false;
            int fails = 0;  // consecutive failed acquire passes
            for (long it = 0; it < 2000000; ++it) {
                if (mine > 0) {
                    // run a task; sometimes publish half of the rest (the others may take them)
                    if (mine > 1 && rng() % 4 == 0) {
                        const int give = mine / 2;
                        mine -= give;
                        pool.fetch_add(give, std::memory_order_release);
                    }
                    mine -= 1;
                    outstanding.fetch_sub(1);
                    continue;
                }
                // acquire: take from the pool, leaving the idle count first
                int p = pool.load(std::memory_order_relaxed);
                if (p > 0) {
                    if (registered) {
                        idle.fetch_sub(1);
                        registered = false;
                    }
                    if (pool.compare_exchange_strong(p, p - 1, std::memory_order_acquire)) {
                        mine = 1;
                        fails = 0;
                        continue;
                    }
                    continue;
                }
                if (!registered && fails++ >= register_after) {
                    registered = true;
                    const int c = idle.fetch_add(1) + 1;
                    if (c == started.load()) {
                        if (outstanding.load() != 0) early.store(true);
                        done.store(1);
                    }
                }
                if (done.load()) break;
                std::this_thread::yield();
            }
        });
    }
    for (auto& t : th) t.join();
    return early.load();
}

}  // namespace quiesce_model

// ------------------------------------------------------------------ quiescence with the rings
namespace ovfq_model {

constexpr int kGroups = 4;

// Workers run tasks; a busy worker sometimes hands half of what it holds to the ring (ticket
// fetch-add, each cell made ready, then every group's epoch bumped, as ovf_spill_deque does).
// An idle worker looks at the ring when its group's epoch (worker w reads group w % kGroups)
// moved since it last found the ring empty, taking a ready head task (leaving the idle count
// first), and records the epoch when it finds the ring empty. With `epoch_rule` it counts
// itself idle only when the epoch has not moved; without it, on its first idle try, and it
// looks at the ring only on every fourth. Returns true if done was set while tasks were
// outstanding.
bool run(unsigned seed, int workers, int work, bool epoch_rule) {
    const uint32_t cap = 1u << 16;
    std::vector<std::atomic<uint8_t>> ready(cap);  // cell p holds a task (no laps here)
    for (auto& r : ready) r.store(0);
    std::atomic<uint64_t> head{0}, tickets{0};
    std::atomic<uint32_t> epoch[kGroups];
    for (auto& e : epoch) e.store(0);
    std::atomic<int> idle{0}, started{1}, done{0}, outstanding{work};
    std::atomic<bool> early{false};
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w)
        th.emplace_back([&, w] {
            std::mt19937 rng(seed * 13 + w);
            if (w != 0) started.fetch_add(1);  // worker 0 owns the job, counted at the reset
            int mine = w == 0 ? work : 0;
            bool registered = false;
            uint32_t seen = epoch[w % kGroups].load();
            uint32_t idle_it = 0;
            for (long it = 0; it < 4000000; ++it) {
                if (mine > 0) {
                    idle_it = 0;
                    if (mine > 2 && rng() % 8 == 0) {
                        const int give = mine / 2;
                        mine -= give;
                        const uint64_t base = tickets.fetch_add((uint64_t)give);
                        for (int j = 0; j < give; ++j) {
                            if (rng() % 64 == 0) std::this_thread::yield();  // a cell made ready late
                            ready[(base + j) & (cap - 1)].store(1, std::memory_order_release);
                        }
                        for (auto& e : epoch) e.fetch_add(1);
                    }
                    mine -= 1;
                    outstanding.fetch_sub(1);
                    continue;
                }
                const uint32_t ep = epoch[w % kGroups].load();
                bool seen_empty = ep == seen;
                const bool look = epoch_rule ? ep != seen : (idle_it % 4 == 1);
                ++idle_it;
                if (look) {
                    const uint64_t h = head.load(), tk = tickets.load();
                    if (tk <= h) {
                        seen = ep;
                        seen_empty = true;
                    } else if (ready[h & (cap - 1)].load(std::memory_order_acquire)) {
                        if (registered) {
                            idle.fetch_sub(1);
                            registered = false;
                        }
                        uint64_t expected = h;
                        if (head.compare_exchange_strong(expected, h + 1, std::memory_order_acq_rel)) {
                            ready[h & (cap - 1)].store(0);
                            mine = 1;
                        }
                        continue;
                    }
                }
                if (!registered && (!epoch_rule || seen_empty)) {
                    registered = true;
                    if (idle.fetch_add(1) + 1 == started.load()) {
                        if (outstanding.load() != 0) early.store(true);
                        done.store(1);
                    }
                }
                if (done.load()) break;
                std::this_thread::yield();
            }
        });
    for (auto& t : th) t.join();
    return early.load();
}

}  // namespace ovfq_model

// ------------------------------------------------------------------ additive join
namespace join_model {

constexpr uint64_t kCountMask = 0xFFF;
constexpr int64_t kMaxValue = (int64_t)1 << 35;

void run(unsigned seed, int children, uint32_t queue, int threads) {
    std::mt19937_64 rng(seed);
    std::vector<int64_t> v(children);
    int64_t expect = 0;
    for (int i = 0; i < children; ++i) {
        // the bounds: |value| < 2^35, half of them at the extremes
        const int64_t mag = (i % 2) ? (kMaxValue - 1) : (int64_t)(rng() % (uint64_t)kMaxValue);
        v[i] = (rng() & 1) ? -mag : mag;
        expect += v[i];
    }
    std::atomic<uint64_t> word{(uint64_t)children | ((uint64_t)queue << 12)};  // sum 0
    std::atomic<int> last{0};
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&] {
            for (int i = next.fetch_add(1); i < children; i = next.fetch_add(1)) {
                const uint64_t delta = ((uint64_t)v[i] << 16) - 1u;
                const uint64_t old = word.fetch_add(delta, std::memory_order_acq_rel);
                if ((old & kCountMask) == 1) {
                    last.fetch_add(1);
                    CHECK(((old >> 12) & 0xF) == queue, "join: resume queue lost (%u)", (unsigned)((old >> 12) & 0xF));
                }
                CHECK((old & kCountMask) >= 1, "join: count below zero");
            }
        });
    for (auto& t : th) t.join();
    const uint64_t w = word.load();
    CHECK(last.load() == 1, "join: %d children saw the last count", last.load());
    CHECK((w & kCountMask) == 0 && ((w >> 12) & 0xF) == queue, "join: final count/queue %llx", (unsigned long long)w);
    const int64_t sum = (int64_t)w >> 16;
    CHECK(sum == expect, "join: sum %lld, expected %lld", (long long)sum, (long long)expect);
}

}  // namespace join_model

int main(int argc, char** argv) {
    const unsigned seeds = argc > 1 ? (unsigned)atoi(argv[1]) : 4;
    for (unsigned s = 1; s <= seeds; ++s) {
        deque_model::run(s, 3, 20000 + 1000 * s);
        ring_model::run(s, 3, 3, 200, 64);
        join_model::run(s, 4095, s % 16, 4);
        join_model::run(s + 100, 1 + (int)(s * 37 % 4095), 15, 3);
        const bool early = quiesce_model::run(s, 6, 1, 2000, true, 0);
        CHECK(!early, "quiesce: done while tasks were outstanding (seed %u)", s);
        const bool early_late = quiesce_model::run(s, 6, 2, 2000, true, 20000);
        CHECK(!early_late, "quiesce: done while a late job owner's tasks were outstanding (seed %u)", s);
        // the idle count joined after 4 failed passes, as thread_scheduler does
        CHECK(!quiesce_model::run(s, 6, 1, 2000, true, 0, 4), "quiesce, delayed count: early end (seed %u)", s);
        CHECK(!quiesce_model::run(s, 6, 2, 2000, true, 20000, 4),
              "quiesce, delayed count: early end with a late job owner (seed %u)", s);
    }
    for (unsigned s = 1; s <= seeds; ++s) {
        ovf_model::run(s, 3, 3, 3000, 64);
        ovf_model::run(s + 50, 2, 4, 2000, 256);
        CHECK(!ovfq_model::run(s, 5, 3000, true), "ovfq: done while tasks were outstanding (seed %u)", s);
    }
    // without the epoch rule a worker can count itself idle before it looks at the ring, and
    // the last worker to go idle ends the launch early with tasks still in the ring
    bool ring_rule_needed = false;
    for (unsigned s = 200; s < 240 && !ring_rule_needed; ++s) ring_rule_needed = ovfq_model::run(s, 5, 3000, false);
    CHECK(ring_rule_needed, "ovfq: no schedule ended early without the ring rule");
    // the former rule: a job owner that starts after the others went idle is not counted,
    // so they declare quiescence with its whole job outstanding
    const bool old_rule_early = quiesce_model::run(99, 6, 1, 2000, false, 50000);
    CHECK(old_rule_early, "quiesce: the schedule that defeats the former rule did not end early");
    if (g_failures) {
        fprintf(stderr, "%d failures\n", g_failures);
        return 1;
    }
    printf("protocol model ok\n");
    return 0;
}
