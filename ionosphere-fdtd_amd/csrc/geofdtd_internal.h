// Declarations shared by the translation units of libgeofdtd.so.
#pragma once

// Records `msg` as the calling thread's last error (geo_last_error) and
// returns `code`.
int geo_internal_fail(int code, const char* msg);
