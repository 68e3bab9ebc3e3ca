// main.cpp -- `pliscuda <conf.txt>`: the CLI of PLISCuda/src/main.cpp:8-70 on
// the MI355X library.  Exit status: 0 ok, 1 usage / no device, 2 parse
// error (the reference prints the message and exits 0), 3-4 run errors.
// `pliscuda --parse-only <conf.txt>` prints the parsed sizes and needs no GPU.
#include <cstring>
#include <fstream>
#include <iostream>

#include "driver.hpp"
#include "plis.h"

int main(int argc, char* argv[]) {
    bool parse_only = argc >= 2 && std::strcmp(argv[1], "--parse-only") == 0;
    const char* conf_path = parse_only ? (argc >= 3 ? argv[2] : nullptr) : (argc >= 2 ? argv[1] : nullptr);
    if (!conf_path) {
        std::cout << "Usage: ./pliscuda [--parse-only] <configuration file>\n";
        return 1;
    }
    std::cout << "\n    **   PLISCuda (MI355X)  **    \n\n";
    if (!parse_only) {
        char name[256];
        const int nd = plis_device_count(name, sizeof name);
        if (nd < 1) {
            std::cout << "No HIP-capable device detected.\n";
            return 1;
        }
        std::cout << "-- HIP device detected -- \n\n  Device name: " << name << "\n\n";
    }
    std::cout << "-- Parsing input files -- \n\n";
    std::ifstream cf(conf_path);
    if (!cf.good()) {
        std::cout << "Couldn't load configuration file\n";
        return 1;
    }
    try {
        std::cout << "  processing configuration file...\n";
        PLISCuda::Configuration conf(cf);
        if (parse_only) {
            const auto& t = conf.getTopology();
            long long tot = 0;
            for (auto c : conf.getState().counts()) tot += c;
            std::cout << "parsed N=" << t.getN() << " S=" << conf.getReactions().getS()
                      << " R=" << conf.getReactions().getR() << " C=" << conf.getCompartments()
                      << " E=" << t.col().size() << " molecules=" << tot << " endTime=" << conf.getEndTime()
                      << " logFreq=" << conf.getLogFreq() << "\n";
            return 0;
        }
        return PLISCuda::run_simulation(conf.getTopology(), conf.getState(), conf.getReactions(),
                                        conf.getReactionsConstantsArray(), conf.getDiffusionConstantsArray(),
                                        conf.getEndTime(), conf.getCompartments(), conf.getCompArray(),
                                        conf.getLogFreq(), conf.options());
    } catch (std::exception& e) {
        std::cout << "  parsing of configuration file failed with msg:\n";
        std::cout << "\t" << e.what() << "\n";
        return 2;
    }
}
