// Online autotuner of the extension kernels' launch shape (items per stolen tile and
// persistent workgroups per CU).
//
// Launches are grouped into buckets by (kernel mode, single-pass last level, parent pattern
// size).  Within a bucket, every candidate shape is tried once on a launch large enough
// to time (HIP events around the launch, read at the next host synchronisation); after
// that the shape with the least time per work item is used.  The shapes never change
// results, only speed: tile size trades tile steals against LDS use, and so against the
// workgroups resident per CU.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "gs_internal.h"

class LaunchTuner {
   public:
    static constexpr long long kMinRows = 4096;  // smaller launches are not timed for tuning

    // shapes per kernel mode and parent size class (the specialised kernels differ in
    // registers, so in the workgroups that fit a CU)
    void init(int cus, bool enabled, int forced_tile) {
        enabled_ = enabled && forced_tile == 0;
        for (int f = 0; f < GS_ITEMS_MODES; ++f) {
            for (int pc = 0; pc < GS_PM_CLASSES; ++pc) {
                auto& cs = cfgs_[f][pc];
                const int m = pc == 2 ? 3 : pc + 1;  // a parent size of the class
                cs.clear();
                for (int tile : {512, 1024, 256}) {  // the first shape is the default
                    if (forced_tile && tile != forced_tile) continue;
                    const int occ = gsk::items_blocks_per_cu(f, tile, m);
                    for (int blocks : {occ, occ / 2}) {
                        if (blocks < 1) continue;
                        bool dup = false;
                        for (auto& c : cs) dup |= c.tile == tile && c.grid == blocks * cus;
                        if (!dup) cs.push_back(gsk::ItemsLaunch{tile, blocks * cus, m});
                        if (!enabled_) break;  // fixed shape: full occupancy only
                    }
                }
            }
        }
    }

    // Shape for a launch over R parent rows in a kernel mode (GS_ITEMS_*); *bucket / *cfg
    // identify it for record() (-1 when the launch is not part of the search).
    gsk::ItemsLaunch choose(int mode, bool single, int m, long long R, int* bucket, int* cfg) {
        const auto& cs = cfgs_[mode][gsk::items_pm_class(m)];
        *bucket = -1;
        *cfg = 0;
        auto with_m = [m](gsk::ItemsLaunch lc) {
            lc.m = m;
            return lc;
        };
        if (!enabled_) return with_m(cs[0]);
        const int b = mode * 1000 + (single ? 100 : 0) + m;
        Bucket& bk = buckets_[b];
        if (bk.ms.empty()) bk.resize(cs.size());
        int best = best_of(bk);
        if (R < kMinRows) {
            *cfg = best;
            return with_m(cs[best]);
        }
        for (size_t i = 0; i < cs.size(); ++i)
            if (bk.n[i] == 0) {
                best = (int)i;
                break;
            }
        *bucket = b;
        *cfg = best;
        return with_m(cs[best]);
    }

    void record(int bucket, int cfg, double ms, long long items) {
        if (bucket < 0 || items <= 0) return;
        Bucket& bk = buckets_[bucket];
        bk.ms[cfg] += ms;
        bk.items[cfg] += (double)items;
        bk.n[cfg] += 1;
    }

    // {"fill single m=3": {"best": [tile, grid], "ns_per_item": [[tile, grid, ns], ...]}, ...}
    std::string report() const {
        std::string s = "{";
        bool first = true;
        for (const auto& kv : buckets_) {
            const int b = kv.first;
            const Bucket& bk = kv.second;
            const int mode = b / 1000;
            const auto& cs = cfgs_[mode][gsk::items_pm_class(b % 100)];
            if (!first) s += ", ";
            first = false;
            static const char* const names[GS_ITEMS_MODES] = {"count", "sets", "rows", "newv"};
            s += "\"" + std::string(names[mode]) + ((b % 1000) >= 100 ? " single" : "") +
                 " m=" + std::to_string(b % 100) + "\": ";
            const int best = best_of(bk);
            s += "{\"best\": [" + std::to_string(cs[best].tile) + ", " +
                 std::to_string(cs[best].grid) + "], \"ns_per_item\": [";
            for (size_t i = 0; i < cs.size(); ++i) {
                if (i) s += ", ";
                const double ns = bk.n[i] ? bk.ms[i] * 1e6 / bk.items[i] : -1.0;
                s += "[" + std::to_string(cs[i].tile) + ", " + std::to_string(cs[i].grid) + ", " +
                     std::to_string(ns) + "]";
            }
            s += "]}";
        }
        return s + "}";
    }

   private:
    struct Bucket {
        std::vector<double> ms, items;
        std::vector<int> n;
        void resize(size_t k) {
            ms.assign(k, 0.0);
            items.assign(k, 0.0);
            n.assign(k, 0);
        }
    };
    static int best_of(const Bucket& bk) {
        int best = 0;
        double bv = 0;
        bool have = false;
        for (size_t i = 0; i < bk.n.size(); ++i) {
            if (!bk.n[i] || bk.items[i] <= 0) continue;
            const double v = bk.ms[i] / bk.items[i];
            if (!have || v < bv) {
                best = (int)i;
                bv = v;
                have = true;
            }
        }
        return best;
    }

    bool enabled_ = false;
    std::vector<gsk::ItemsLaunch> cfgs_[GS_ITEMS_MODES][GS_PM_CLASSES];
    std::map<int, Bucket> buckets_;
};
