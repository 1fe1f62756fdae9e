/* Enumerations and constants of the LibRPA C API.
 * Values and names match LibRPA include/librpa_enums.h:11-111 so that callers
 * compiled against LibRPA's headers link against this library unchanged. */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

#define LIBRPA_UNSET -101
#define LIBRPA_AUTO -51
#define LIBRPA_SWITCH_OFF 0
#define LIBRPA_SWITCH_ON 1
#define LIBRPA_VERBOSE_DEBUG 4
#define LIBRPA_VERBOSE_WARN 3
#define LIBRPA_VERBOSE_INFO 2
#define LIBRPA_VERBOSE_CRITICAL 1
#define LIBRPA_VERBOSE_SILENT 0

#define LIBRPA_ROUTING_COUNT 5
/* Parallel routing (librpa_enums.h:58-64).  On one GPU every routing runs the
 * same device-resident pipeline; the value is accepted and recorded. */
typedef enum {
    ROUTING_UNSET = LIBRPA_UNSET,
    AUTO = LIBRPA_AUTO,
    RTAU = 0,
    ATOMPAIR = 1,
    LIBRI = 2,
} LibrpaParallelRouting;

#define LIBRPA_TFGRID_COUNT 7
/* Time/frequency grid type (librpa_enums.h:76-84). */
typedef enum {
    TFGRID_UNSET = LIBRPA_UNSET,
    GaussLegendre = 0,
    GaussChebyshevI = 1,
    GaussChebyshevII = 2,
    Minimax = 3,
    EvenSpaced = 4,
    EvenSpaced_TF = 5,
} LibrpaTimeFreqGrid;

typedef int LibrpaSwitch;
typedef int LibrpaKind;
typedef int LibrpaVerbose;

#ifdef __cplusplus
}
#endif
