// Cross-lane synchronisation of a run enqueued item by item on `nl` streams
// (lanes). Host code only, no HIP.
//
// Item i runs on lane[i] after the earlier items deps[i]. Items of one lane
// run in enqueue order. An item waits (hipStreamWaitEvent) for a dependency
// on another lane only if its lane does not already know that dependency
// complete: each lane keeps a vector clock -- per lane, the last position
// known complete -- raised by its own items and by the clocks of the items
// it waited for (an item's clock is its lane's clock when it was enqueued:
// everything the item's lane had waited for is complete before the item
// starts). Of several dependencies on one lane only the latest is awaited.
// An item records its event only if some later item waits for it.
#pragma once

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace kunrt {

struct LaneSync {
    std::vector<std::vector<int>> waits; // per item: items of other lanes to wait for
    std::vector<char> record;            // per item: some item waits for it
};

inline LaneSync plan_lane_sync(const std::vector<int> &lane, const std::vector<std::vector<int>> &deps, int nl) {
    const size_t n = lane.size();
    LaneSync p;
    p.waits.resize(n);
    p.record.assign(n, 0);
    std::vector<std::vector<int>> clock(nl, std::vector<int>(nl, -1)); // per lane: positions known complete
    std::vector<int> next(nl, 0), pos(n), best(nl);
    std::vector<std::vector<int>> snap(n); // per item: its lane's clock once it is enqueued
    for (size_t i = 0; i < n; i++) {
        const int X = lane[i];
        if (X < 0 || X >= nl) throw std::runtime_error("lane out of range");
        std::vector<int> &c = clock[X];
        std::fill(best.begin(), best.end(), -1);
        for (int d : deps[i]) {
            if (d < 0 || (size_t)d >= i) throw std::runtime_error("launch DAG out of topological order");
            const int Y = lane[d];
            if (Y != X && (best[Y] < 0 || pos[d] > pos[best[Y]])) best[Y] = d;
        }
        for (int Y = 0; Y < nl; Y++) {
            const int d = best[Y];
            if (d < 0 || pos[d] <= c[Y]) continue; // already known complete
            p.waits[i].push_back(d);
            p.record[d] = 1;
            for (int Z = 0; Z < nl; Z++) c[Z] = std::max(c[Z], snap[d][Z]);
        }
        pos[i] = next[X]++;
        c[X] = pos[i];
        snap[i] = c;
    }
    return p;
}

} // namespace kunrt
