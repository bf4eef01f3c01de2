/*
 * readparams.h -- drop-in for PSBA/readparams.h:8-11.
 *
 * Same signature and meaning as the reference's reader; errors print a
 * message and terminate, as the reference does (PSBA/readparams.cpp:463-524).
 * Implemented on psba_read_sba (psba_io.h).
 */
#pragma once

#include "psba.h"

#define FULLQUATSZ 4

void readInitialSBAEstimate(char* camsfname, char* ptsfname, int cnp, int pnp, int mnp,
                            void (*infilter)(dtype* pin, int nin, dtype* pout, int nout), int cnfp, int* ncams,
                            int* n3Dpts, int* n2Dprojs, dtype** motstruct, dtype** initrot, dtype** imgpts,
                            dtype** covimgpts, char** vmask);
