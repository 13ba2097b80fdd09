// Wall-clock Timer and leveled printing with the names and arguments of
// MRCPP's utilities (src/utils/Timer.h:40-58, src/utils/Printer.h:59-133), so
// programs written against MRCPP's API compile unchanged.  A message prints
// when its level is at most the level given to Printer::init; the layout of
// the lines is this library's own.
#pragma once

#include <chrono>
#include <complex>
#include <cstdio>
#include <iostream>
#include <string>

#include "function_tree.h"

namespace mrcpp {

class Timer final {
public:
    explicit Timer(bool start_timer = true) {
        if (start_timer) start();
    }
    void start() {
        elapsed_ = 0.0;
        resume();
    }
    void resume() {
        running_ = true;
        t0_ = std::chrono::steady_clock::now();
    }
    void stop() {
        if (!running_) return;
        elapsed_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        running_ = false;
    }
    // seconds, including the running interval
    double elapsed() const {
        return running_ ? elapsed_ + std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count() : elapsed_;
    }

private:
    bool running_ = false;
    double elapsed_ = 0.0;
    std::chrono::steady_clock::time_point t0_{};
};

class Printer final {
public:
    static void init(int level = 0, int rank = 0, int size = 1, const char * /*file*/ = nullptr) {
        state().level = rank == 0 ? level : -1;
        (void)size;
    }
    static int setPrintLevel(int i) {
        const int old = state().level;
        state().level = i;
        return old;
    }
    static int getPrintLevel() { return state().level; }
    static int setPrecision(int i) {
        const int old = state().prec;
        state().prec = i;
        return old;
    }
    static int getPrecision() { return state().prec; }
    static int setWidth(int i) {
        const int old = state().width;
        state().width = i;
        return old;
    }
    static int getWidth() { return state().width; }
    static bool enabled(int level) { return level <= state().level; }

private:
    struct State {
        int level = -1, prec = 5, width = 70;
    };
    static State &state() {
        static State s;
        return s;
    }
};

namespace print {
inline void separator(int level, const char &c, int newlines = 0) {
    if (!Printer::enabled(level)) return;
    std::cout << std::string(Printer::getWidth(), c) << std::string(newlines + 1, '\n');
}
inline void environment(int level) {
    if (!Printer::enabled(level)) return;
    separator(level, '-');
    std::cout << " MRCPP-compatible operator application on AMD Instinct (gfx950), FP64 drop-in\n";
    separator(level, '-', 1);
}
inline void header(int level, const std::string &txt, int newlines = 0, const char &c = '=') {
    if (!Printer::enabled(level)) return;
    separator(level, c);
    std::cout << " " << txt << "\n";
    separator(level, '-', newlines);
}
inline void value(int level, const std::string &txt, double v, const std::string &unit = "", int p = -1, bool sci = true) {
    if (!Printer::enabled(level)) return;
    char buf[64];
    std::snprintf(buf, sizeof(buf), sci ? "%.*e" : "%.*f", p < 0 ? Printer::getPrecision() : p, v);
    std::printf(" %-32s %-8s %s\n", txt.c_str(), unit.c_str(), buf);
}
inline void complex_value(int level, const std::string &txt, std::complex<double> v, const std::string &unit = "", int p = -1, bool sci = true) {
    value(level, txt + " (re)", v.real(), unit, p, sci);
    value(level, txt + " (im)", v.imag(), unit, p, sci);
}
inline void time(int level, const std::string &txt, const Timer &timer) { value(level, txt, timer.elapsed(), "(sec)", 5, false); }
inline void tree(int level, const std::string &txt, int n, int m, double t) {
    if (!Printer::enabled(level)) return;
    std::printf(" %-24s %8d nodes %8d end nodes %10.5f sec\n", txt.c_str(), n, m, t);
}
inline void tree(int level, const std::string &txt, const FunctionTree<3> &tree, const Timer &timer) {
    print::tree(level, txt, tree.getNNodes(), tree.getNEndNodes(), timer.elapsed());
}
// host memory is not tracked per call here
inline void memory(int level, const std::string &txt) {
    if (!Printer::enabled(level)) return;
    std::printf(" %-32s\n", txt.c_str());
}
inline void footer(int level, const Timer &timer, int newlines = 0, const char &c = '=') {
    if (!Printer::enabled(level)) return;
    separator(level, '-');
    std::printf(" Wall time: %.5f sec\n", timer.elapsed());
    separator(level, c, newlines);
}
} // namespace print

} // namespace mrcpp
