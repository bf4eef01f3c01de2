/*
 * psba.h -- constants and return codes of the drop-in interface
 * (same names and values as PSBA/psba.h:3-18).
 */
#pragma once

#include <stdio.h>

#define K_DIM 5
#define dtype double

#define PSBA_INIT_MU 1E-03
#define PSBA_STOP_THRESH 1e-12
#define PSBA_EPSILON 1E-12
#define PSBA_EPSILON2 1E-12
#define PSBA_EPSILON_SQ ((PSBA_EPSILON) * (PSBA_EPSILON))

#define ITER_TURN_TO_LM 1
#define ITER_TURN_TO_TR 2
#define ITER_CONTINUE 3
#define ITER_ERR 4
#define ITER_DP_NO_CHANGE 5
#define ITER_ERR_SMALL_ENOUGH 6
#define ITER_PASS 7
