// Random launch DAGs through plan_lane_sync (kunquant_amd/csrc/lane_sync.hpp):
// every dependency of every item must be complete before the item starts,
// following only stream order and the planned waits, and every awaited item
// must record its event. Prints the waits planned against one wait per
// cross-lane dependency; exits 1 on the first violation.
#include <cstdio>
#include <random>
#include <vector>

#include "lane_sync.hpp"

int main() {
    std::mt19937 rng(7);
    long planned = 0, naive = 0;
    for (int trial = 0; trial < 400; trial++) {
        const int n = 1 + (int)(rng() % 400), nl = 1 + (int)(rng() % 16);
        std::vector<int> lane(n);
        std::vector<std::vector<int>> deps(n);
        for (int i = 0; i < n; i++) {
            lane[i] = (int)(rng() % nl);
            const int k = i ? (int)(rng() % 5) : 0;
            for (int q = 0; q < k; q++) deps[i].push_back((int)(rng() % i));
            for (int d : deps[i]) naive += lane[d] != lane[i];
        }
        const kunrt::LaneSync p = kunrt::plan_lane_sync(lane, deps, nl);
        // known[i]: items complete before item i starts
        const int W = (n + 63) / 64;
        std::vector<std::vector<unsigned long long>> known(n, std::vector<unsigned long long>(W, 0));
        std::vector<int> last(nl, -1);
        auto add = [&](std::vector<unsigned long long> &k, int d) {
            for (int w = 0; w < W; w++) k[w] |= known[d][w];
            k[d / 64] |= 1ull << (d % 64);
        };
        for (int i = 0; i < n; i++) {
            std::vector<unsigned long long> k(W, 0);
            if (last[lane[i]] >= 0) add(k, last[lane[i]]);
            for (int d : p.waits[i]) {
                if (d >= i || lane[d] == lane[i] || !p.record[d]) {
                    std::printf("trial %d item %d: bad wait on %d\n", trial, i, d);
                    return 1;
                }
                add(k, d);
                planned++;
            }
            for (int d : deps[i])
                if (!((k[d / 64] >> (d % 64)) & 1)) {
                    std::printf("trial %d item %d: dependency %d not complete\n", trial, i, d);
                    return 1;
                }
            known[i] = k;
            last[lane[i]] = i;
        }
    }
    std::printf("ok: %ld waits planned for %ld cross-lane dependencies\n", planned, naive);
    return 0;
}
