/* Handler management (LibRPA include/librpa_handler.h:25-57).
 *
 * LibRPA's header includes <mpi.h> for the communicator argument.  This
 * single-GPU library runs as one rank; it accepts the communicator and
 * ignores it.  Define LIBRPA_USE_MPI_H to take MPI_Comm from <mpi.h>;
 * otherwise a single-rank stand-in with MPICH's integer handle type is
 * declared (ABI-compatible with MPICH callers). */
#pragma once

#if defined(LIBRPA_USE_MPI_H)
#include <mpi.h>
#elif !defined(MPI_COMM_WORLD)
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque handle: only the instance id is visible (librpa_handler.h:26-31). */
typedef struct {
    const int instance_id_;
} LibrpaHandler;

/* Create an instance; replaces librpa_create_handler (librpa_handler.h:45). */
LibrpaHandler *librpa_create_handler(MPI_Comm comm);

/* Free an instance and all its device data; NULL is ignored (librpa_handler.h:57). */
void librpa_destroy_handler(LibrpaHandler *h);

/* Entry point of the Fortran binding (binding/fortran/librpa_f03.f90:407, implemented in
 * src/api/handler.cpp:26-30): Fortran communicator handle by reference. */
LibrpaHandler *librpa_create_handler_fortran(int *f_comm);

#ifdef __cplusplus
}
#endif
