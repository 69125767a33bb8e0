/*
 * op_profile.h — the timing-tree C API of the MI355X-native OP2 backend
 * (libop2hip.so).  Applications include it by this name, as the reference's
 * applications do (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:58); it carries
 * the reference's C entry points (op2/include/op_profile.h:197-212) without
 * its C++ classes.
 *
 * Expected use: op_profile_start(name), then nested op_profile_enter(name) /
 * op_profile_exit() (or op_profile_next(name)) around application sections,
 * op_profile_end(), op_profile_output().
 *
 * OP_PROFILE_LEVEL (read at op_profile_start): 0 off, 1 (default) the
 * application's sections, 2 also one node per op_par_loop
 * ("<loop>/hip/<plan>"), 3 also the loop's computation and the kernels'
 * device time (HIP events).  OP_PROFILE_JSON_OUTPUT=<file>: op_profile_output
 * also writes the tree as JSON.  Timed section boundaries synchronise the
 * device, except inside a HIP graph capture.
 */
#ifndef OP2_HIP_OP_PROFILE_H
#define OP2_HIP_OP_PROFILE_H

#ifdef __cplusplus
extern "C" {
#endif

void op_profile_start(const char *name);
void op_profile_enter(const char *name);
void op_profile_enter_kernel(const char *name, const char *target,
                             const char *variant);
void op_profile_next(const char *name);
void op_profile_exit(void);
void op_profile_end(void);
void op_profile_output(void);
void op_profile_output_json(const char *filename);

#ifdef __cplusplus
}
#endif

#endif
