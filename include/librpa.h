/* Umbrella header (LibRPA include/librpa.h:1-31). */
#pragma once

#include "librpa_global.h"
#include "librpa_input.h"
#include "librpa_compute.h"
#include "librpa_amd.h"
