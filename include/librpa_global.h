/* Global functions (LibRPA include/librpa_global.h:26-88). */
#pragma once

#include "librpa_enums.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Build description string (librpa_global.h:26). */
const char *librpa_get_build_info(void);
/* Version components (librpa_global.h:33-47). */
int librpa_get_major_version(void);
int librpa_get_minor_version(void);
int librpa_get_patch_version(void);

/* Initialise the library: selects the GPU and sets up logging
 * (librpa_global.h:62-64; C++ default arguments as in LibRPA). */
#ifdef __cplusplus
void librpa_init_global(LibrpaSwitch switch_redirect_stdout = LIBRPA_SWITCH_OFF,
                        const char *redirect_path = "stdout",
                        LibrpaSwitch switch_process_output = LIBRPA_SWITCH_ON);
#else
void librpa_init_global(LibrpaSwitch switch_redirect_stdout, const char *redirect_path,
                        LibrpaSwitch switch_process_output);
#endif

/* Entry point of the Fortran binding (binding/fortran/librpa_f03.f90:316, implemented in
 * src/api/global.cpp:56-61). */
void librpa_init_global_fortran(int *f_comm, LibrpaSwitch switch_redirect_stdout, const char *redirect_path,
                                LibrpaSwitch switch_process_output);

/* Release every instance (librpa_global.h:72). */
void librpa_finalize_global(void);

/* Self test: runs a small device contraction (librpa_global.h:80). */
void librpa_test(void);

/* Print the per-region timing table (librpa_global.h:88). */
void librpa_print_profile(void);

#ifdef __cplusplus
}
#endif
