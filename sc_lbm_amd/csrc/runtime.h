// runtime.h -- library state shared by the translation units of libsclbm.so.
#pragma once

namespace scl {

// Initialises the library's device (once; checks that it is a gfx950) and
// returns SCL_OK or an error code with the message recorded for
// scl_last_error().
int runtime_ready();
// The HIP device the library runs on.
int runtime_device();

}  // namespace scl
