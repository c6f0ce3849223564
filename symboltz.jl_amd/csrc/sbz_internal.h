// Library-internal helpers: status codes and the thread-local last-error message.
#pragma once
#include "../../include/sbz.h"

void sbz_set_error(const char* fmt, ...);
void sbz_clear_error();
