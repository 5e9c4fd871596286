// lbm_slots.hpp -- the slot table behind the asynchronous snapshots
// (lbm_snapshot / lbm_snapshot_wait / lbm_snapshot_release).
//
// Plain C++ (no HIP), so its state machine is unit-tested on the CPU under
// ThreadSanitizer.  A slot goes
//
//   FREE --take--> ENQUEUED --waited(ok)--> HELD --release--> FREE
//                           --waited(failed)--> FAILED --release--> FREE
//                  ENQUEUED --give_back (the enqueue itself failed)--> FREE
//
// take() blocks while every slot is taken; once any wait has failed it
// returns at once with an error instead, because the failed copy's slot may
// never come back (its caller may have stopped) and a blocked producer would
// hang.  release() accepts HELD and FAILED slots, so a consumer whose wait
// failed still hands the slot back.  All methods may be called from any thread.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <mutex>

namespace lbm_host {

template <int NSLOTS>
class SlotTable {
   public:
    enum State { FREE = 0, ENQUEUED = 1, HELD = 2, FAILED = 3 };
    enum Result { OK = 0, BAD_TICKET = 1, NOT_WAITED = 2, EARLIER_FAILURE = 3 };

    // Reserve a free slot for a new ticket: on OK, *slot and *ticket are set and
    // the slot is ENQUEUED.  Blocks while none is free; EARLIER_FAILURE once a
    // wait has failed.
    Result take(int* slot, int64_t* ticket) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (failed_) return EARLIER_FAILURE;
            for (int i = 0; i < NSLOTS; ++i)
                if (state_[i] == FREE) {
                    state_[i] = ENQUEUED;
                    ticket_[i] = next_++;
                    *slot = i;
                    *ticket = ticket_[i];
                    return OK;
                }
            cv_.wait(lk);
        }
    }

    // The enqueue behind a take() failed: the slot is free again.
    void give_back(int slot) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            state_[slot] = FREE;
            ticket_[slot] = -1;
        }
        cv_.notify_all();
    }

    // Slot of a ticket that is not free, or -1.
    int find(int64_t ticket) const {
        std::lock_guard<std::mutex> lk(mu_);
        for (int i = 0; i < NSLOTS; ++i)
            if (state_[i] != FREE && ticket_[i] == ticket) return i;
        return -1;
    }

    // The wait for `slot` ended: HELD if it succeeded, else FAILED (and every
    // later take() fails at once).
    void waited(int slot, bool ok) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (ok) {
                state_[slot] = HELD;
            } else {
                state_[slot] = FAILED;
                failed_ = true;
            }
        }
        cv_.notify_all();
    }

    // Hand a waited (HELD or FAILED) slot back.
    Result release(int64_t ticket) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            int s = -1;
            for (int i = 0; i < NSLOTS; ++i)
                if (state_[i] != FREE && ticket_[i] == ticket) s = i;
            if (s < 0) return BAD_TICKET;
            if (state_[s] == ENQUEUED) return NOT_WAITED;
            state_[s] = FREE;
            ticket_[s] = -1;
        }
        cv_.notify_all();
        return OK;
    }

    State state(int slot) const {
        std::lock_guard<std::mutex> lk(mu_);
        return State(state_[slot]);
    }

    bool failed() const {
        std::lock_guard<std::mutex> lk(mu_);
        return failed_;
    }

   private:
    mutable std::mutex mu_;
    std::condition_variable cv_;
    int state_[NSLOTS] = {};
    int64_t ticket_[NSLOTS] = {};
    int64_t next_ = 1;
    bool failed_ = false;
};

}  // namespace lbm_host
