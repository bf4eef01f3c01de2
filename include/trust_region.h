/*
 * trust_region.h -- dogleg / 2-D subspace driver, same signature as
 * PSBA/trust_region.h:5-11 (blk_idx is accepted and unused, as in the
 * reference).  Uses and advances the global `int itno`.
 */
#pragma once

#include "cl_psba.h"
#include "psba.h"

int trust_region(PSBA_structPtr psbaPtr, int cnp, int pnp, int mnp, int n3Dpts, int nCams, int n2Dprojs,
                 int* blk_idx, dtype* finalErr);
