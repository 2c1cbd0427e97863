// bmx_io.h -- internal: error hand-off from the I/O module to bmx_last_error.
#pragma once

extern "C" void bmx_io_set_error(const char *msg);
