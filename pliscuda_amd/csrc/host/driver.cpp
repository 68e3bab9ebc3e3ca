// driver.cpp -- the host loop of `pliscuda conf.txt` on top of libplis_hip.so.
//
// Mirrors PLISCuda::run_simulation (PLISCuda/src/driver.cu:12-408): the same
// stdout report (final simulation time, elapsed time, simulation steps,
// steps/second) and the same snapshot files sim<timestamp>_<t>.dat written
// every logFreq simulated seconds and at the end.  All stepping is done on
// the device by plis_run_until; the host only pauses at snapshot times (a
// pause never shortens a step, so snapshots do not perturb the trajectory).
#include "driver.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <iterator>
#include <limits>
#include <string>
#include <vector>

#include "plis.h"

namespace PLISCuda {

void print_eltime(float secs) {
    const int hrs = (int)(secs / 3600);
    const int mts = (int)((secs - hrs * 3600) / 60);
    const float scs = secs - (hrs * 3600 + mts * 60);
    std::cout << hrs << "h " << mts << "m ";
    std::printf("%5.2fs", scs);
    std::fflush(stdout);
}

std::string print_state_snapshot(const uint32_t* h_state, int spc, unsigned sbc, float current_time) {
    std::string s;
    s += "subvolumes=" + std::to_string(sbc) + "\n";
    s += "species=" + std::to_string(spc) + "\n";
    s += "time=" + std::to_string(current_time) + "\n\n";
    for (unsigned i = 0; i < sbc; ++i) {
        for (int j = 0; j < spc; ++j) s += std::to_string(h_state[(size_t)j * sbc + i]) + " ";
        s += "\n";
    }
    return s;
}

// The final state in the input State format ("subvolumes: N", "species: S",
// then "v: c0 c1 ..."), which the reference's snapshots are not.
static bool write_state_file(const std::string& path, const std::vector<uint32_t>& x, int S, unsigned N) {
    std::ofstream f(path);
    if (!f) return false;
    f << "subvolumes: " << N << "\nspecies: " << S << "\n\n";
    for (unsigned v = 0; v < N; ++v) {
        f << v << ":";
        for (int s = 0; s < S; ++s) f << " " << x[(size_t)s * N + v];
        f << "\n";
    }
    return (bool)f;
}

static void write_snapshot(std::time_t tstamp, int real, int nreal, const std::vector<uint32_t>& x, int S,
                           unsigned N, double t) {
    const float tf = (float)t;
    std::string name = "sim" + std::to_string(tstamp) + "_";
    if (nreal > 1) name += "r" + std::to_string(real) + "_";
    name += std::to_string(tf) + ".dat";
    std::ofstream f(name);
    f << print_state_snapshot(x.data(), S, N, tf);
}

int run_simulation(Topology t, State s, Reactions r, float* h_rrc, float* h_drc, float endTime,
                   int compartments_count, int* subv_constants, float log_freq, const Options& opts) {
    const unsigned N = t.getN();
    const int S = s.getS(), R = r.getR(), C = compartments_count;
    std::cout << "\n-- Starting Simulation -- \n\n";
    std::vector<double> k((size_t)R * C), d((size_t)S * C);
    for (size_t i = 0; i < k.size(); ++i) k[i] = (double)h_rrc[i];
    for (size_t i = 0; i < d.size(); ++i) d[i] = (double)h_drc[i];
    plis_network net{S, R, C, r.reactants().data(), r.products().data(), k.data(), d.data()};
    plis_topology topo{N, (uint32_t)t.col().size(), t.rowPtr().data(), t.col().data(), nullptr, subv_constants};
    const int nreal = opts.realizations < 1 ? 1 : opts.realizations;
    std::vector<plis_topology> topos(nreal, topo);
    std::vector<const uint32_t*> xs(nreal, s.counts().data());
    uint64_t seed = opts.seed;
    if (seed == 0) seed = (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
    plis_params params{opts.epsilon, opts.n_critical, 0, 0, 1, seed};
    plis_ctx* ctx = nullptr;
    int rc = plis_create(&net, nreal, topos.data(), xs.data(), nullptr, &params, &ctx);
    if (rc != PLIS_OK) {
        std::cerr << "plis_create failed: " << plis_last_error(nullptr) << "\n";
        return 3;
    }
    if (!opts.restore.empty()) {
        std::ifstream in(opts.restore, std::ios::binary);
        std::vector<char> blob((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        if (!in.good() && !in.eof()) blob.clear();
        if (blob.empty() || plis_restore(ctx, blob.data(), blob.size()) != PLIS_OK) {
            std::cerr << "cannot restore from " << opts.restore << ": "
                      << (blob.empty() ? "unreadable file" : plis_last_error(ctx)) << "\n";
            plis_destroy(ctx);
            return 3;
        }
        std::cout << "  [restored from " << opts.restore << "]\n";
    }
    int32_t shape[3];
    plis_launch_shape(ctx, shape);
    std::cout << "  [using " << shape[1] << " block(s) of " << shape[0] << " wavefront(s), " << shape[2]
              << " tiles of 64 subvolumes]\n";
    std::cout << "  [epsilon " << opts.epsilon << ", n_c " << opts.n_critical << ", seed " << seed << ", "
              << nreal << " realization(s)]\n\n";
    std::cout << "-- Begin Iterations -- \n\n";
    std::cout.precision(5);

    const std::time_t tstamp = std::time(nullptr);
    const auto wall0 = std::chrono::steady_clock::now();
    const double t_end = (double)endTime;
    const uint64_t no_limit = std::numeric_limits<uint64_t>::max();
    std::vector<uint32_t> x((size_t)S * N);
    // per realization: the clock of its last snapshot (realizations step
    // with their own tau, so they reach their snapshot times at different
    // global iterations)
    std::vector<double> last_log(nreal, 0.0);
    if (!opts.restore.empty()) {
        for (int q = 0; q < nreal; ++q) {
            plis_stats st;
            plis_get_stats(ctx, q, &st);
            last_log[q] = st.t;  // snapshots continue from the restored clock
        }
    }
    bool ok = true;
    if (log_freq > 0) {
        for (;;) {
            // pause at the earliest snapshot time still due; a realization
            // already past it does not step in this call
            double t_pause = std::numeric_limits<double>::infinity();
            for (int q = 0; q < nreal; ++q) {
                plis_stats sq;
                plis_get_stats(ctx, q, &sq);
                if (sq.status == 0 && sq.t < t_end) t_pause = std::min(t_pause, last_log[q] + (double)log_freq);
            }
            if (t_pause == std::numeric_limits<double>::infinity()) break;  // every realization finished
            rc = plis_run_until(ctx, t_end, t_pause, no_limit);
            if (rc != PLIS_OK) break;
            for (int q = 0; q < nreal; ++q) {
                plis_stats sq;
                plis_get_stats(ctx, q, &sq);
                // a realization paused on crossing its own snapshot time
                // (finished ones get their final snapshot below)
                if (sq.status != 0 || sq.t >= t_end || sq.t - last_log[q] < (double)log_freq) continue;
                last_log[q] = sq.t;
                plis_get_state(ctx, q, x.data());
                write_snapshot(tstamp, q, nreal, x, S, N, sq.t);
            }
        }
    } else {
        rc = plis_run(ctx, t_end, no_limit);
    }
    if (rc != PLIS_OK) {
        std::cerr << "simulation failed: " << plis_last_error(ctx) << "\n";
        ok = false;
    }
    plis_sync(ctx);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    plis_stats st0;
    plis_get_stats(ctx, 0, &st0);
    if (st0.status == 2) std::printf("\n -- WARNING: min(tau) = +Inf or < 0 - abort simulation --\n");
    std::cout << "\n-- Simulation Complete -- \n\n";
    uint64_t steps = 0, leapbacks = 0, firings = 0, leaps = 0;
    for (int q = 0; q < nreal; ++q) {
        plis_stats sq;
        plis_get_stats(ctx, q, &sq);
        plis_get_state(ctx, q, x.data());
        write_snapshot(tstamp, q, nreal, x, S, N, sq.t);
        if (!opts.state_out.empty()) {
            const std::string path = nreal > 1 ? opts.state_out + "_r" + std::to_string(q) : opts.state_out;
            if (!write_state_file(path, x, S, N)) {
                std::cerr << "cannot write " << path << "\n";
                ok = false;
            }
        }
        steps += sq.steps;
        leapbacks += sq.leapbacks;
        firings += sq.firings;
        leaps += sq.subvolume_leaps;
    }
    std::cout << "  final simulation time:  " << st0.t << "s\n";
    std::cout << "  elapsed time:           ";
    print_eltime((float)el);
    std::cout << "\n  simulation steps:       " << steps << "\n";
    std::cout << "  steps/second:           " << (el > 0 ? steps / el : 0.0) << "\n";
    std::cout << "  leap-backs:             " << leapbacks << "\n";
    std::cout << "  firings:                " << firings << "\n";
    std::cout << "  subvolume-leaps/second: " << (el > 0 ? leaps / el : 0.0) << "\n";
    std::cout << "\n";
    if (!opts.checkpoint.empty()) {
        uint64_t bytes = 0;
        std::vector<char> blob;
        bool wrote = plis_checkpoint_size(ctx, &bytes) == PLIS_OK;
        if (wrote) {
            blob.resize(bytes);
            wrote = plis_checkpoint(ctx, blob.data(), bytes) == PLIS_OK;
        }
        if (wrote) {
            std::ofstream f(opts.checkpoint, std::ios::binary);
            f.write(blob.data(), (std::streamsize)blob.size());
            wrote = (bool)f;
        }
        if (!wrote) {
            std::cerr << "cannot write checkpoint " << opts.checkpoint << "\n";
            ok = false;
        }
    }
    plis_destroy(ctx);
    return ok ? 0 : 4;
}

}  // namespace PLISCuda
