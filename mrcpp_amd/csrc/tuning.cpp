// Autotuner of the band sum's launch shape: the wave layout of the order-7
// matrix-core kernel, the minimum groups per work unit and the persistent
// grid, chosen by timing the caller's own workload.  No reference
// counterpart: MRCPP's OpenMP node loop has no launch shape to tune.
#include <algorithm>
#include <chrono>
#include <vector>

#include "device.h"
#include "function_tree.h"

namespace mrcpp {

namespace {

struct LaunchState {
    int layout, unitGroups, wgPerCu;
};

LaunchState current() {
    LaunchState s{};
    s.layout = mwg_band_layout();
    mwg_band_launch(&s.unitGroups, &s.wgPerCu);
    return s;
}

void set(const LaunchState &s) {
    device::check(mwg_set_band_layout(s.layout), "band layout");
    device::check(mwg_set_band_launch(s.unitGroups, s.wgPerCu), "band launch");
}

} // namespace

BandTuning autotune_band(const std::function<void()> &run, int repeats, std::vector<BandTuning> *timings) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    device::ensure_init();
    if (!run) throw Error("autotune: no workload");
    if (repeats < 1) throw Error("autotune: repeats must be positive");
    const LaunchState before = current();
    std::vector<BandTuning> all;
    try {
        for (int layout : {MWG_BAND_ROWS, MWG_BAND_TILES})
            for (int ug : {24, 32, 48})
                for (int wg : {0, 3}) {
                    set({layout, ug, wg});
                    run(); // warm-up: operand tables, scratch sizes
                    std::vector<double> t;
                    for (int r = 0; r < repeats; r++) {
                        device::check(mwg_synchronize(), "synchronize");
                        const auto t0 = std::chrono::steady_clock::now();
                        run();
                        device::check(mwg_synchronize(), "synchronize");
                        t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
                    }
                    std::sort(t.begin(), t.end());
                    all.push_back({layout, ug, wg, t[t.size() / 2]});
                }
    } catch (...) {
        set(before);
        throw;
    }
    const BandTuning best = *std::min_element(all.begin(), all.end(), [](const BandTuning &a, const BandTuning &b) { return a.seconds < b.seconds; });
    set({best.layout, best.unitGroups, best.wgPerCu});
    if (timings) *timings = all;
    return best;
}

} // namespace mrcpp
