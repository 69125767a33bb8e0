// op2h_profile.h — what the loop lifecycle tells the timing tree of
// op2h_profile.cpp (the op_profile_* C API).
#pragma once

namespace op2h {

// True between op_profile_start and op_profile_end at OP_PROFILE_LEVEL >= 2
// when the current section is timed: op_par_loops then get nodes, and the
// loop lifecycle synchronises the device around each loop to time it.
bool profile_loops_on();

// The level read at op_profile_start (1 before it).
int profile_level();

// One op_par_loop call: a loop node "<name>/hip/<variant>" under the current
// section gets `wall_ns`; at level 3 its "Computation" child gets `comp_ns`
// and that child's "Kernel" child the kernels' device time `kernel_ms` (none
// when negative).
void profile_loop(const char *name, const char *variant, long long wall_ns,
                  long long comp_ns, double kernel_ms);

}  // namespace op2h
