// Runs a set of cases in one GPU context: all active cases advance together through lbm_step (one
// launch per kernel family and step, whatever their box sizes), and the host is called back only at
// the steps where some case needs a measurement or a field file.  A case that stops (its last step,
// or an unstable state) is left out of later launches.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "itolbm.h"

namespace itolbm_host {

struct Member {
    virtual ~Member() = default;
    virtual lbm_case_desc desc() const = 0;
    virtual int steps() const = 0;                                // n_steps of the case
    virtual bool init(lbm_ctx* ctx, int id) = 0;                  // initial condition
    virtual int next_event(int s) const = 0;                      // first step >= s the host must see
    // called after step t (post-collision state of step t on the device); false = the case stops at t
    virtual bool on_event(lbm_ctx* ctx, int id, int t) = 0;
    // steps completed (n_steps, or the failing step) and the wall time since the batch started
    virtual void finish(int steps_done, double elapsed) = 0;
    virtual const std::string& name() const = 0;
};

// Returns 0, or 1 if the context could not be created or a library call failed.
inline int run_batch(std::vector<Member*>& members, int precision)
{
    if (members.empty()) return 0;
    std::vector<lbm_case_desc> descs;
    for (auto* m : members) descs.push_back(m->desc());
    lbm_ctx* ctx = nullptr;
    if (lbm_create(descs.data(), (int)descs.size(), precision, &ctx) != LBM_OK) {
        std::fprintf(stderr, "lbm_create failed: %s\n", lbm_last_error(ctx));
        lbm_destroy(ctx);
        return 1;
    }
    auto fail = [&](const char* what) {
        std::fprintf(stderr, "%s: %s\n", what, lbm_last_error(ctx));
        lbm_destroy(ctx);
        return 1;
    };
    const int n = (int)members.size();
    std::vector<char> active(n, 1);
    for (int i = 0; i < n; ++i)
        if (!members[i]->init(ctx, i)) return fail("initial condition");
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    auto stop = [&](int i, int steps_done) -> int {
        active[i] = 0;
        if (lbm_sync(ctx) != LBM_OK) return 1;
        members[i]->finish(steps_done, elapsed());
        return lbm_set_active(ctx, i, 0);
    };
    for (int i = 0; i < n; ++i)
        if (members[i]->steps() <= 0 && stop(i, 0)) return fail("finish");
    int s = 0;   // next step to execute, common to every active case
    for (;;) {
        int e = -1;
        for (int i = 0; i < n; ++i)
            if (active[i]) e = (e < 0) ? members[i]->next_event(s) : std::min(e, members[i]->next_event(s));
        if (e < 0) break;
        if (lbm_step(ctx, e - s + 1, LBM_STEP_RECORD_AUX) != LBM_OK) return fail("lbm_step");
        for (int i = 0; i < n; ++i) {
            if (!active[i] || members[i]->next_event(s) != e) continue;
            if (!members[i]->on_event(ctx, i, e) && stop(i, e)) return fail("finish");
        }
        s = e + 1;
        for (int i = 0; i < n; ++i)
            if (active[i] && s >= members[i]->steps() && stop(i, members[i]->steps())) return fail("finish");
    }
    lbm_destroy(ctx);
    return 0;
}

}  // namespace itolbm_host
