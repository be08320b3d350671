// Shared host-side helpers of libgsim: thread-local error text and status
// plumbing used by every C-ABI entry point.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "gsim.h"

namespace gsim {

// Formats msg into the thread-local error slot read by gsim_last_error() and
// returns st, so call sites can write `return fail(GSIM_ERROR_..., "...")`.
gsim_status fail(gsim_status st, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));

// Clears the thread-local error slot.
void clear_error();

}  // namespace gsim
