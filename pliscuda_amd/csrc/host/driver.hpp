// driver.hpp -- PLISCuda::run_simulation over the C ABI of libplis_hip.so.
//
// Same signature and argument meaning as the reference
// (PLISCuda/include/cuda/driver.cuh:32-33): topology, initial state, reactions,
// reaction constants [c*R + j], diffusion constants [c*S + s], end time,
// compartment count, compartment id per subvolume, snapshot interval.
// Differences (deliberate): returns a status (0 ok, non-zero error) instead of
// exit()ing inside; the Options argument carries epsilon / n_c / seed /
// realizations, which the reference hard-codes.
#pragma once

#include "plis_inputs.hpp"

namespace PLISCuda {

int run_simulation(Topology t, State s, Reactions r, float* h_rrc, float* h_drc, float endTime,
                   int compartments_count, int* subv_constants, float log_freq,
                   const Options& opts = Options());

// snapshot text (PLISCuda/src/driver.cu:435-447)
std::string print_state_snapshot(const uint32_t* h_state, int spc, unsigned sbc, float current_time);

void print_eltime(float secs);

}  // namespace PLISCuda
