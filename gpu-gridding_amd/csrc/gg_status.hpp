// gg_status.hpp -- the library's per-thread status (mi355_grid_last_error /
// _message), shared by its translation units.
#pragma once

namespace ggs {

// Records status and a printf-style message for the calling thread; returns
// status.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// Records success; returns MI355_OK.
int ok();

} // namespace ggs
