// plis_inputs.hpp -- PLISCuda's input surface for the C++ driver.
//
// Same file grammars, key order and error messages as the reference parsers
// (PLISCuda/include/cpp/Configuration.hpp:27-144, Topology.hpp:98-191,
// State.hpp:112-217, Reactions.hpp:132-223, cpp_utils.hpp:29-56), written
// from the format description; errors throw std::invalid_argument like the
// reference.  Topology keeps any number of neighbours per subvolume (CSR);
// the reference stores 6 slots padded with the subvolume's own id.
//
// Optional keys may follow the reference's keys in conf.txt (the reference
// hard-codes them, driver.cu:36-37): epsilon=, nc=, seed=, realizations=.
#pragma once

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace PLISCuda {

class Topology {
   public:
    unsigned int getN() const { return n_; }
    const std::vector<uint32_t>& rowPtr() const { return rp_; }
    const std::vector<uint32_t>& col() const { return col_; }
    void read(std::istream& is);

   private:
    unsigned int n_ = 0;
    std::vector<uint32_t> rp_, col_;
};

class State {
   public:
    unsigned int getN() const { return n_; }
    int getS() const { return s_; }
    // species-major [s*N + v] like the reference's getState()
    const std::vector<uint32_t>& counts() const { return x_; }
    void read(std::istream& is);

   private:
    unsigned int n_ = 0;
    int s_ = 0;
    std::vector<uint32_t> x_;
};

class Reactions {
   public:
    int getR() const { return r_; }
    int getS() const { return s_; }
    // [s*R + j] like getReactants() / getProducts()
    const std::vector<int32_t>& reactants() const { return reac_; }
    const std::vector<int32_t>& products() const { return prod_; }
    void read(std::istream& is);

   private:
    int r_ = 0, s_ = 0;
    std::vector<int32_t> reac_, prod_;
};

struct Options {
    double epsilon = 0.05;  // reference value (driver.cu:37)
    int n_critical = 10;    // reference value (driver.cu:36)
    uint64_t seed = 0;      // 0: derived from the wall clock
    int realizations = 1;
    std::string checkpoint;  // write a checkpoint here at the end of the run
    std::string restore;     // continue from this checkpoint
    std::string state_out;   // final state in the input State format (re-readable)
};

class Configuration {
   public:
    explicit Configuration(std::ifstream& config_file);

    Topology& getTopology() { return topology_; }
    Reactions& getReactions() { return reactions_; }
    State& getState() { return state_; }
    float getEndTime() const { return end_time_; }
    float getLogFreq() const { return log_freq_; }
    int getCompartments() const { return compartments_; }
    int* getCompArray() { return comp_.data(); }
    float* getReactionsConstantsArray() { return k_.data(); }
    float* getDiffusionConstantsArray() { return d_.data(); }
    const Options& options() const { return opts_; }

   private:
    Topology topology_;
    Reactions reactions_;
    State state_;
    float end_time_ = 0.f, log_freq_ = 0.f;
    int compartments_ = 1;
    std::vector<int> comp_;
    std::vector<float> k_, d_;
    Options opts_;
};

}  // namespace PLISCuda
