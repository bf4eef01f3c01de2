/*
 * levmar.h -- Levenberg-Marquardt driver, same signature as PSBA/levmar.h:6-11.
 * Uses and advances the caller-defined global `int itno` (PSBA/main.cpp:27).
 */
#pragma once

#include "cl_psba.h"
#include "psba.h"

int levmar(PSBA_structPtr psbaPtr, int cnp, int pnp, int mnp, int n3Dpts, int nCams, int n2Dprojs,
           dtype* finalErr);

/* Sum of squared residuals at the entry of the last levmar call
 * (PSBA/levmar.cpp:16,95).  The application may define it, as PSBA/main.cpp:24
 * does; the library's weak definition is used otherwise. */
extern dtype initErr;
