// parse_check.cpp -- test harness: parses one configuration (and the files it
// names) with the host input layer, the way `pliscuda --parse-only` does,
// without linking the GPU library, so it can be built with AddressSanitizer
// and UBSan on the host.  Exit status 0 parsed, 2 rejected with a message,
// 1 configuration file unreadable.
#include <fstream>
#include <iostream>

#include "plis_inputs.hpp"

int main(int argc, char* argv[]) {
    if (argc < 2) return 1;
    std::ifstream cf(argv[1]);
    if (!cf.good()) return 1;
    try {
        PLISCuda::Configuration conf(cf);
        long long tot = 0;
        for (auto c : conf.getState().counts()) tot += c;
        std::cout << "parsed N=" << conf.getTopology().getN() << " S=" << conf.getReactions().getS()
                  << " R=" << conf.getReactions().getR() << " C=" << conf.getCompartments()
                  << " E=" << conf.getTopology().col().size() << " molecules=" << tot << "\n";
        return 0;
    } catch (std::exception& e) {
        std::cout << "rejected: " << e.what() << "\n";
        return 2;
    }
}
