// slots_test.cpp -- the snapshot slot state machine (lbmcl_amd/csrc/lbm_slots.hpp)
// under concurrent producers and consumers, built with and without sanitizers
// (tests/test_sanitizers.py runs it under ThreadSanitizer and under
// AddressSanitizer + UndefinedBehaviorSanitizer).
//
// Checks:
//   * tickets are unique and increasing; a slot is reused only after release;
//   * take() blocks while every slot is held and resumes on release;
//   * releasing an enqueued (not yet waited) or unknown ticket is an error;
//   * a failed wait leaves its slot FAILED: release() still frees it, and every
//     later take() returns EARLIER_FAILURE at once -- also one that was
//     already blocked when the failure happened (no hang);
//   * a producer/consumer pipeline over many tickets on several threads.
// Exit status 0 when every check passes.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <deque>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "lbm_slots.hpp"

using Table = lbm_host::SlotTable<3>;

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s\n", what);
        ++failures;
    }
}

void basic_transitions() {
    Table t;
    int s[3];
    int64_t k[3];
    for (int i = 0; i < 3; ++i) check(t.take(&s[i], &k[i]) == Table::OK, "take a free slot");
    check(k[0] == 1 && k[1] == 2 && k[2] == 3, "tickets count up from 1");
    check(s[0] != s[1] && s[1] != s[2] && s[0] != s[2], "distinct slots");
    check(t.release(k[0]) == Table::NOT_WAITED, "release before wait is refused");
    check(t.release(99) == Table::BAD_TICKET, "unknown ticket");
    check(t.find(k[1]) == s[1] && t.find(99) == -1, "find");
    t.waited(s[0], true);
    check(t.state(s[0]) == Table::HELD, "waited -> HELD");
    check(t.release(k[0]) == Table::OK, "release a held slot");
    check(t.release(k[0]) == Table::BAD_TICKET, "double release");
    int s4;
    int64_t k4;
    check(t.take(&s4, &k4) == Table::OK && s4 == s[0] && k4 == 4, "a released slot is reused with a new ticket");
    t.give_back(s4);
    check(t.state(s4) == Table::FREE && t.find(k4) == -1, "give_back frees the slot");
}

void blocking_take_resumes_on_release() {
    Table t;
    int s;
    int64_t k[3];
    for (int i = 0; i < 3; ++i) t.take(&s, &k[i]);
    for (int i = 0; i < 3; ++i) t.waited(t.find(k[i]), true);
    std::atomic<bool> got{false};
    std::thread producer([&] {
        int s2;
        int64_t k2;
        const bool ok = t.take(&s2, &k2) == Table::OK;
        got = ok;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    check(!got.load(), "take blocks while every slot is held");
    check(t.release(k[1]) == Table::OK, "release from another thread");
    producer.join();
    check(got.load(), "the blocked take resumes after a release");
}

void failed_wait_does_not_hang() {
    Table t;
    int s;
    int64_t k[3];
    for (int i = 0; i < 3; ++i) t.take(&s, &k[i]);
    t.waited(t.find(k[0]), true);
    t.waited(t.find(k[1]), true);
    std::atomic<int> result{-1};
    std::thread producer([&] {  // blocks: every slot is taken
        int s2;
        int64_t k2;
        result = int(t.take(&s2, &k2));
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    t.waited(t.find(k[2]), false);  // the copy of k[2] failed
    producer.join();
    check(result.load() == Table::EARLIER_FAILURE, "a blocked take returns once a wait failed");
    check(t.failed(), "the table records the failure");
    check(t.state(t.find(k[2])) == Table::FAILED, "the failed slot is FAILED");
    check(t.release(k[2]) == Table::OK, "a failed slot can be released");
    int s3;
    int64_t k3;
    check(t.take(&s3, &k3) == Table::EARLIER_FAILURE, "takes fail at once after a failure, even with free slots");
    check(t.release(k[0]) == Table::OK && t.release(k[1]) == Table::OK, "held slots are still released");
}

void pipeline() {
    // two producers take tickets, one consumer waits and releases them in order
    Table t;
    std::mutex mu;
    std::deque<int64_t> queue;
    std::set<int64_t> seen;
    std::atomic<int> produced{0};
    const int per_producer = 2000;
    auto produce = [&] {
        for (int i = 0; i < per_producer; ++i) {
            int s;
            int64_t k;
            if (t.take(&s, &k) != Table::OK) return;
            std::lock_guard<std::mutex> lk(mu);
            queue.push_back(k);
            ++produced;
        }
    };
    bool ok = true;
    std::thread consumer([&] {
        int done = 0;
        while (done < 2 * per_producer) {
            int64_t k = -1;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!queue.empty()) {
                    k = queue.front();
                    queue.pop_front();
                }
            }
            if (k < 0) {
                std::this_thread::yield();
                continue;
            }
            const int s = t.find(k);
            if (s < 0 || t.state(s) != Table::ENQUEUED) ok = false;
            t.waited(s, true);
            if (t.release(k) != Table::OK) ok = false;
            if (!seen.insert(k).second) ok = false;
            ++done;
        }
    });
    std::thread p1(produce), p2(produce);
    p1.join();
    p2.join();
    consumer.join();
    check(ok && produced.load() == 2 * per_producer && int(seen.size()) == 2 * per_producer,
          "4000 tickets through 3 slots: each taken, waited and released once");
}

}  // namespace

int main() {
    basic_transitions();
    blocking_take_resumes_on_release();
    failed_wait_does_not_hang();
    pipeline();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("slot table: all checks passed\n");
    return 0;
}
