/*
 * The library's own stubs for the five loops of the airfoil application
 * (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:225-277), with their
 * hand-written gfx950 kernels: an application written against this library
 * may call them instead of translating its loops.  Their names are the ones
 * the translator gives those loops (op_par_loop_<stem>_<i>_<kernel>,
 * cpp/loop_host.hpp.jinja:35-41), so this header is kept apart from
 * op2_hip.h: a translated program defines its own stubs under the same names,
 * possibly with other argument lists (the rms-only update of
 * airfoil_tempdats/dp and airfoil_plain/sp takes five arguments), and must
 * not see these declarations.  Within libop2hip.so calls bind to the
 * library's own definitions (linked -Bsymbolic-functions), so a program's
 * stubs of the same name do not replace them.
 */
#ifndef OP2_HIP_AIRFOIL_H
#define OP2_HIP_AIRFOIL_H

#include "op2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Argument lists exactly as airfoil.cpp:225-277 passes them. */
void op_par_loop_airfoil_1_save_soln(char const *name, op_set set, op_arg arg0,
                                     op_arg arg1);
void op_par_loop_airfoil_2_adt_calc(char const *name, op_set set, op_arg arg0,
                                    op_arg arg1, op_arg arg2, op_arg arg3,
                                    op_arg arg4, op_arg arg5);
void op_par_loop_airfoil_3_res_calc(char const *name, op_set set, op_arg arg0,
                                    op_arg arg1, op_arg arg2, op_arg arg3,
                                    op_arg arg4, op_arg arg5, op_arg arg6,
                                    op_arg arg7);
void op_par_loop_airfoil_4_bres_calc(char const *name, op_set set, op_arg arg0,
                                     op_arg arg1, op_arg arg2, op_arg arg3,
                                     op_arg arg4, op_arg arg5);
void op_par_loop_airfoil_5_update(char const *name, op_set set, op_arg arg0,
                                  op_arg arg1, op_arg arg2, op_arg arg3,
                                  op_arg arg4, op_arg arg5, op_arg arg6,
                                  op_arg arg7);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* OP2_HIP_AIRFOIL_H */
