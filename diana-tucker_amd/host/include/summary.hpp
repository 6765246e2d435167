// Summary profiler with the reference's static interface (include/summary.hpp,
// src/summary.cpp:19-206): nested named regions with flop and byte counts,
// one table on stderr at the end.  Regions are timed on the host clock (every
// host-layer call returns after its GPU work, like the reference's MPI_Wtime
// regions around synchronous operators); while recording, the library also
// times every kernel launch with HIP events on its stream, and the table is
// followed by those device times per kernel.  Flop and byte counts are filled
// by the host layer (Function::, HOOI_ALS sweeps, Sp_HOOI_ALS sweeps) with the
// algorithmic figures of the work it runs, so GFlop/s and GB/s are populated.
#ifndef DIANA_AMD_SUMMARY_HPP
#define DIANA_AMD_SUMMARY_HPP
#include <cassert>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "communicator.hpp"
#include "device.hpp"

inline std::string method_name(const std::string &pretty) {
    const size_t colons = pretty.find("::");
    const size_t begin = pretty.substr(0, colons).rfind(' ') + 1;
    const size_t end = pretty.rfind('(') - begin;
    return pretty.substr(begin, end) + "()";
}
#define METHOD_NAME method_name(__PRETTY_FUNCTION__)

class Summary {
    struct Event {
        std::string name;
        size_t caller;
        long long flop, bytes;
        double t0, t1, length, counted;  // seconds; counted = time inside callees
    };
    static constexpr size_t kRoot = (size_t)-1;
    static inline std::vector<Event> events_;
    static inline std::map<std::string, std::vector<size_t>> by_name_;
    static inline size_t current_ = kRoot;
    static inline bool recording_ = false;

    static double now() {
        using clk = std::chrono::steady_clock;
        return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
    }
    static std::string pad(std::string s, size_t len) {
        if (s.size() < len) s.append(len - s.size(), ' ');
        return s;
    }
    static std::string rule(int a, int b, int c) {
        return "|-" + std::string(a, '-') + "+-" + std::string(b, '-') + "+-" + std::string(c, '-') + "| \n";
    }

public:
    // Starts recording and opens the "{Main}" region; device kernel timing on.
    static void init() {
        events_.clear();
        by_name_.clear();
        current_ = kRoot;
        recording_ = true;
        dt_timing_reset(diana::Device::ctx());
        dt_timing_enable(diana::Device::ctx(), 1);
        start("{Main}");
    }
    static void finalize() {
        end("{Main}");
        recording_ = false;
        dt_timing_enable(diana::Device::ctx(), 0);
    }
    static bool recording() { return recording_; }

    // Opens a region inside the current one; flop and bytes are this region's
    // own work (callees add theirs to their callers when they end).
    static void start(const std::string &name, long long flop = 0, long long bytes = 0) {
        if (!recording_) return;
        events_.push_back({name, current_, flop, bytes, now(), 0.0, 0.0, 0.0});
        current_ = events_.size() - 1;
        by_name_[name].push_back(current_);
    }
    static void end(const std::string &name) {
        if (!recording_) return;
        assert(current_ != kRoot && events_[current_].name == name);
        (void)name;
        Event &e = events_[current_];
        e.t1 = now();
        e.length = e.t1 - e.t0;
        current_ = e.caller;
        if (current_ != kRoot) {
            Event &c = events_[current_];
            c.flop += e.flop;
            c.bytes += e.bytes;
            c.counted += e.length;
        }
    }

    // The reference's table: per region name its total time, the share of it
    // spent in nested regions ("Kernel" when none), call count, average time,
    // GFlop/s (1e9 flop) and bandwidth (2^30 bytes per second); then the
    // library's per-kernel device times of the recorded span.
    static void print_summary() {
        const int cap = 15, first = 10;
        const std::string sep = "| ";
        std::string out = rule(first, 4 * cap, 2 * cap);
        out += sep + pad("", first) + sep;
        for (const char *c : {"Time(s)", "Time C.(%)", "Number", "Avg. Time(s)"}) out += pad(c, cap);
        out += sep;
        for (const char *c : {"GFlop/s", "Bandw.(GB/s)"}) out += pad(c, cap);
        out += sep + "\n" + rule(first, 4 * cap, 2 * cap);
        for (const auto &kv : by_name_) {
            double total = 0.0, counted = 0.0;
            long long flop = 0, bytes = 0;
            for (size_t i : kv.second) {
                total += events_[i].length;
                counted += events_[i].counted;
                flop += events_[i].flop;
                bytes += events_[i].bytes;
            }
            const double gf = total > 0 ? (double)flop / 1e9 / total : 0.0;
            const double gb = total > 0 ? (double)bytes / 1073741824.0 / total : 0.0;
            out += sep + pad(kv.first, first + 4 * cap + sep.size()) + sep;
            out += (kv.first == "{Main}") ? pad(std::to_string(gf), cap) + pad(std::to_string(gb), cap)
                                          : pad("", 2 * cap);
            out += sep + "\n" + sep + pad("", first) + sep;
            out += pad(std::to_string(total), cap);
            out += pad(counted == 0.0 ? std::string("Kernel") : std::to_string(counted / total * 100.0), cap);
            out += pad(std::to_string(kv.second.size()), cap);
            out += pad(std::to_string(total / (double)kv.second.size()), cap);
            out += sep + pad(std::to_string(gf), cap) + pad(std::to_string(gb), cap) + sep + "\n";
            out += rule(first, 4 * cap, 2 * cap);
        }
        std::vector<char> buf(1 << 16);
        if (dt_timing_list(diana::Device::ctx(), buf.data(), buf.size()) == DT_OK && buf[0]) {
            out += "GPU kernels (HIP events on the library stream): name launches total_ms\n";
            out += buf.data();
        }
        if (mpi_rank() == 0) std::cerr << out << std::endl;
    }
};

// Scoped region for the host layer: start() now, end() when it leaves scope.
struct SummaryScope {
    std::string name;
    SummaryScope(std::string n, long long flop = 0, long long bytes = 0) : name(std::move(n)) {
        Summary::start(name, flop, bytes);
    }
    ~SummaryScope() { Summary::end(name); }
};
#endif
